/*
 * amr_core.h -- per-thread logic of the AMR deposit / refinement / relink
 * kernels (gfx950).  Every function here works on one node, one row or one
 * octet of a level and is called by exactly one GPU thread.
 *
 * Level representation.  Every AHF grid level at resolution L is a union of
 * "octets": the 2x2x2 children of the refined nodes of the coarser level
 * (AHF refines in x-runs of even start and even length and mirrors every
 * refined row and plane, src/libamr_serial/refine_grid.c:44-107, 155-896,
 * so no fine node exists without its seven siblings; the domain grid is the
 * union of all octets of a virtual L/2 grid).  Per level:
 *   - octets sorted by (oz, oy, ox) at resolution L/2, packed in oct_key;
 *   - node n = 8*octet + (cz*4 + cy*2 + cx)  ("octet-major" order);
 *   - oct_link bit 0: the x-run (AHF nquad) continues into the octet at ox+1;
 *   - nbr[27*o + (ez+1)*9 + (ey+1)*3 + (ex+1)]: the octet at the periodic
 *     offset (ex,ey,ez) or -1;
 *   - octet rows (oy, oz) and octet planes (oz) as CSR ranges of octets.
 * AHF's pquads are then the maximal runs of consecutive present octet
 * planes, its cquads the maximal runs of consecutive present octet rows of
 * a plane and its nquads the linked runs of octets of a row (two blocks of
 * the same kind never touch except across the periodic wrap).
 *
 * Particle lists.  AHF's node lists are head-inserted: ll() builds them in
 * descending particle index (src/libamr_serial/lltools.c:85-86) and every
 * relink() reverses the order of what it moves (src/libamr_serial/
 * relink.c:163-164, 209-211).  A level's list order is therefore descending
 * particle index on even levels and ascending on odd levels; buckets are
 * stored ascending and walked in the level's direction.
 */
#ifndef AHF_AMR_CORE_H
#define AHF_AMR_CORE_H

#include <stdint.h>
#if !defined(__HIPCC__)
#include <math.h>
#endif

#if defined(__HIPCC__)
#define AHF_HD __host__ __device__ inline
#else
#define AHF_HD static inline
#endif

#define AHF_KEY_BITS 21
#define AHF_KEY_MASK ((1ull << AHF_KEY_BITS) - 1ull)
#define AHF_FULL27 0x7FFFFFFu
#define AHF_MIN_NNODES 125
#define AHF_NO_OCTET (-1)

/* node flags (byte per node) */
#define NF_INTERIOR 1u  /* all 27 TSC neighbours present (test_tsc)        */
#define NF_TEST 2u      /* test_node() true: interior and density criterion */
#define NF_REFINED 4u   /* becomes an octet of the next level               */
#define NF_LINK 8u      /* refined and the fine x-run continues to x+1      */

typedef struct {
  int64_t L;                /* nodes per dimension of this level           */
  int64_t n_oct;
  const uint64_t *oct_key;  /* ox | oy<<21 | oz<<42 (resolution L/2)       */
  const uint8_t *oct_link;  /* bit 0: run continues to ox+1               */
  const int32_t *nbr;       /* 27 per octet                                */
  int64_t n_row;            /* octet rows                                  */
  const uint64_t *row_key;  /* oy | oz<<21                                 */
  const int32_t *row_oct0;  /* [n_row+1] first octet of each row           */
  const int32_t *row_plane; /* octet plane of each row                     */
  int64_t n_plane;
  const int32_t *plane_z;   /* oz of each octet plane                      */
  const int32_t *plane_row0;/* [n_plane+1] first row of each plane         */
} level_view;

/* bucket entry record of dense levels: FP32 position of the entry's
 * particle (w unused) */
typedef struct __attribute__((aligned(16))) {
  float x, y, z, w;
} ahf_q4;

/* TSC weights of one bucket entry against its own node, for the offsets
 * d = -1, 0, +1 (index d+1) of the node it contributes to:
 *   a[k]  = mass2dens * w(sz, k-1)   (the first product of AHF's
 *           mass2dens * wz * wy * wx, density.c:393-400)
 *   wy[k] = w(sy, k-1), wx[k] = w(sx, k-1)                               */
typedef struct __attribute__((aligned(16))) {
  double a[3], wy[3], wx[3], pad;
} ahf_w9;

AHF_HD uint64_t ahf_pack3(uint64_t x, uint64_t y, uint64_t z)
{
  return x | (y << AHF_KEY_BITS) | (z << (2 * AHF_KEY_BITS));
}
AHF_HD int64_t ahf_kx(uint64_t k) { return (int64_t)(k & AHF_KEY_MASK); }
AHF_HD int64_t ahf_ky(uint64_t k) { return (int64_t)((k >> AHF_KEY_BITS) & AHF_KEY_MASK); }
AHF_HD int64_t ahf_kz(uint64_t k) { return (int64_t)((k >> (2 * AHF_KEY_BITS)) & AHF_KEY_MASK); }
/* octet-row keys hold oy | oz<<21 */
AHF_HD int64_t ahf_row_y(uint64_t k) { return (int64_t)(k & AHF_KEY_MASK); }

/* offset index of (ex,ey,ez) in {-1,0,1}^3 */
AHF_HD int ahf_e27(int ex, int ey, int ez) { return (ez + 1) * 9 + (ey + 1) * 3 + (ex + 1); }

/* octet carry of child coordinate c in {0,1} moved by d in {-1,0,1} */
AHF_HD int ahf_carry(int c, int d) { return (c + d) < 0 ? -1 : ((c + d) > 1 ? 1 : 0); }
AHF_HD int ahf_child(int c, int d) { return (c + d) & 1; }

/* node index of the neighbour of node (o,c) at offset d, or -1 */
AHF_HD int64_t ahf_nbr_node(const int32_t *nbr, int64_t o, int c, int dx, int dy, int dz)
{
  int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
  int32_t on = nbr[27 * o + ahf_e27(ahf_carry(cx, dx), ahf_carry(cy, dy), ahf_carry(cz, dz))];
  if (on < 0) return -1;
  return 8 * (int64_t)on + (ahf_child(cz, dz) << 2 | ahf_child(cy, dy) << 1 | ahf_child(cx, dx));
}

/*
 * presence mask of node (o,c): bit (dz+1)*9+(dy+1)*3+(dx+1) set when AHF's
 * get_TSCnodes() finds the neighbour at (dx,dy,dz)
 * (src/libamr_serial/get_nnodes.c:656-853 with search_TSCplane :514-651 and
 * search_TSCcolumn :459-509):
 *   plane z+dz needs the node (x,y,z+dz), column y+dy needs (x,y+dy,z+dz),
 *   and x+dx must lie in the same x-run as (x,y+dy,z+dz) unless the step
 *   crosses the periodic wrap, where the run at the other end only has to
 *   exist.
 */
AHF_HD uint32_t ahf_presence_mask(const level_view *lv, int64_t o, int c)
{
  const int32_t *nb = lv->nbr + 27 * o;
  const int64_t half = lv->L / 2;
  const int64_t ox = ahf_kx(lv->oct_key[o]);
  const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
  uint32_t m = 0;
  for (int dz = -1; dz <= 1; dz++) {
    const int ez = ahf_carry(cz, dz);
    if (nb[ahf_e27(0, 0, ez)] < 0) continue;
    for (int dy = -1; dy <= 1; dy++) {
      const int ey = ahf_carry(cy, dy);
      const int32_t orow = nb[ahf_e27(0, ey, ez)];
      if (orow < 0) continue;
      const int base = (dz + 1) * 9 + (dy + 1) * 3;
      m |= 1u << (base + 1);
      /* x+1 */
      if (cx == 0) m |= 1u << (base + 2);
      else {
        const int32_t ot = nb[ahf_e27(1, ey, ez)];
        if (ot >= 0 && (ox == half - 1 || (lv->oct_link[orow] & 1))) m |= 1u << (base + 2);
      }
      /* x-1 */
      if (cx == 1) m |= 1u << (base + 0);
      else {
        const int32_t ot = nb[ahf_e27(-1, ey, ez)];
        if (ot >= 0 && (ox == 0 || (lv->oct_link[ot] & 1))) m |= 1u << (base + 0);
      }
    }
  }
  return m;
}

/* f1mod, src/libutility/specific.c:120-129 */
AHF_HD double ahf_f1mod(double x)
{
  if (x >= 2.0) return x - 2.0;
  if (x >= 1.0) return x - 1.0;
  return x;
}

/* the three offsets in the order AHF visits the source nodes along one axis
 * (its traversal is raster order of the periodic source coordinate) */
AHF_HD void ahf_axis_order(int64_t X, int64_t L, int ord[3])
{
  if (X == 0) { ord[0] = 0; ord[1] = 1; ord[2] = -1; }
  else if (X == L - 1) { ord[0] = 1; ord[1] = -1; ord[2] = 0; }
  else { ord[0] = -1; ord[1] = 0; ord[2] = 1; }
}

/*
 * One source row of a lane in a staged plane (the march's rows walk): the
 * row's three sources are adjacent nodes, so their entries form the range
 * [lo, hi) of the staged order with inner boundaries b1, b2 (first entries
 * of the second and third source).  AHF's x order of the three is a
 * rotation of the row (ahf_axis_order: rot 0 inside the box, 1 at x = 0,
 * 2 at x = L-1), so the walk starts at the rot-th source and wraps round.
 */
AHF_HD uint32_t ahf_row_start(uint32_t lo, uint32_t b1, uint32_t b2, uint32_t hi, int rot)
{
  const uint32_t q = rot == 0 ? lo : (rot == 1 ? b1 : b2);
  return q == hi ? lo : q; /* the sources from the first to the row's end are empty */
}
AHF_HD uint32_t ahf_row_next(uint32_t q, uint32_t lo, uint32_t hi) { return q + 1u == hi ? lo : q + 1u; }
/* the source of entry q along the row: 0, 1, 2 (ex = -1, 0, +1) */
AHF_HD uint32_t ahf_row_src(uint32_t q, uint32_t b1, uint32_t b2) { return q >= b2 ? 2u : (q >= b1 ? 1u : 0u); }

/* x in [-L, 2L) folded into [0, L) */
AHF_HD int64_t ahf_wrap(int64_t x, int64_t L) { return x < 0 ? x + L : (x >= L ? x - L : x); }

/* weight index k in AHF's weights[k][dim] for target-source = d */
AHF_HD double ahf_tsc_w(double s, int d)
{
  if (d == 0) return 0.75 - s * s;
  if (d < 0) return ((0.5 - s) * (0.5 - s)) / 2;
  return ((0.5 + s) * (0.5 + s)) / 2;
}

/* particle-node separation along one axis, src/libamr_serial/density.c:344-354 */
AHF_HD double ahf_pn_sep(float p, double cen, double L)
{
  double s = ((double)p - cen) * L;
  if (fabs(s) > 0.5 * L) {
    double a = ahf_f1mod(((double)p + 0.5) + 1.0);
    double b = ahf_f1mod((cen + 0.5) + 1.0);
    s = (a - b) * L;
  }
  return s;
}

/* node centre along one axis, src/libamr_serial/density.c:318-324 */
AHF_HD double ahf_centre(int64_t X, double L)
{
  return ahf_f1mod((((double)X) / L) + 0.5 / L + 1.0);
}

/* centres of the nodes at X-1, X, X+1 (periodic) along one axis */
AHF_HD void ahf_centres3(int64_t X, int64_t L, double c[3])
{
  c[0] = ahf_centre(ahf_wrap(X - 1, L), (double)L);
  c[1] = ahf_centre(X, (double)L);
  c[2] = ahf_centre(ahf_wrap(X + 1, L), (double)L);
}

/* contribution mass2dens*wz*wy*wx of a particle with separations (sx,sy,sz)
 * from its node to the node at weight offsets (dx,dy,dz)
 * (src/libamr_serial/density.c:393-400) */
AHF_HD double ahf_contrib(double m, double sx, double sy, double sz, int dx, int dy, int dz)
{
  return m * ahf_tsc_w(sz, dz) * ahf_tsc_w(sy, dy) * ahf_tsc_w(sx, dx);
}

/* coordinates of node (o,c) of a level */
AHF_HD void ahf_node_xyz(const level_view *lv, int64_t o, int c, int64_t *X, int64_t *Y, int64_t *Z)
{
  const uint64_t key = lv->oct_key[o];
  *X = 2 * ahf_kx(key) + (c & 1);
  *Y = 2 * ahf_ky(key) + ((c >> 1) & 1);
  *Z = 2 * ahf_kz(key) + (c >> 2);
}

/* separations of a particle from the centre of the node it is linked to,
 * in grid units, one per axis (src/libamr_serial/density.c:342-355) */
AHF_HD void ahf_particle_sep(const level_view *lv, int64_t node, float x, float y, float z, double sep[3])
{
  const double L = (double)lv->L;
  int64_t X, Y, Z;
  ahf_node_xyz(lv, node >> 3, (int)(node & 7), &X, &Y, &Z);
  sep[0] = ahf_pn_sep(x, ahf_centre(X, L), L);
  sep[1] = ahf_pn_sep(y, ahf_centre(Y, L), L);
  sep[2] = ahf_pn_sep(z, ahf_centre(Z, L), L);
}

/* the weight record of an entry whose particle is at (x,y,z) and is linked
 * to node `node` (separations as in src/libamr_serial/density.c:340-360) */
AHF_HD void ahf_entry_weights(const level_view *lv, int64_t node, float x, float y, float z, double mass2dens,
                              ahf_w9 *w)
{
  double sep[3];
  ahf_particle_sep(lv, node, x, y, z, sep);
  for (int k = 0; k < 3; k++) {
    w->a[k] = mass2dens * ahf_tsc_w(sep[2], k - 1);
    w->wy[k] = ahf_tsc_w(sep[1], k - 1);
    w->wx[k] = ahf_tsc_w(sep[0], k - 1);
  }
  w->pad = 0.0;
}

/* contribution of an entry to the node at offset (-ex,-ey,-ez) from the
 * entry's own node: ((mass2dens*wz)*wy)*wx, AHF's association
 * (src/libamr_serial/density.c:393-400) */
AHF_HD double ahf_w9_contrib(const ahf_w9 *w, int ex, int ey, int ez)
{
  return w->a[1 - ez] * w->wy[1 - ey] * w->wx[1 - ex];
}

/*
 * Exact-order density of one target node from the weight records of the
 * bucket entries, walked as one flat sequence: the sources that contribute
 * (present and non-empty) form a 27-bit mask in traversal order; the walk
 * takes them lowest bit first and each one's entries in list order, and
 * rounds each contribution into the FP32 value as AHF does
 * (src/libamr_serial/density.c:393-400).
 */
AHF_HD float ahf_gather_dens_w(const level_view *lv, const uint32_t *pmask, const int64_t *node_p0,
                               const ahf_w9 *W, int descending, int64_t o, int c)
{
  int64_t X, Y, Z;
  ahf_node_xyz(lv, o, c, &X, &Y, &Z);
  const int rx = X == 0 ? 1 : (X == lv->L - 1 ? 2 : 0);
  const int ry = Y == 0 ? 1 : (Y == lv->L - 1 ? 2 : 0);
  const int rz = Z == 0 ? 1 : (Z == lv->L - 1 ? 2 : 0);
  uint32_t mask = 0;
  int64_t src[27];
  for (int t = 0; t < 27; t++) {
    const int ez = (t / 9 + rz) % 3 - 1, ey = ((t / 3) % 3 + ry) % 3 - 1, ex = (t % 3 + rx) % 3 - 1;
    const int64_t s = ahf_nbr_node(lv->nbr, o, c, ex, ey, ez);
    src[t] = s;
    if (s < 0) continue;
    if (!((pmask[s] >> ((1 - ez) * 9 + (1 - ey) * 3 + (1 - ex))) & 1u)) continue;
    if (node_p0[s + 1] > node_p0[s]) mask |= 1u << t;
  }
  float dens = -1.0f;
  while (mask) {
    int t = 0;
    while (!((mask >> t) & 1u)) t++;
    mask &= mask - 1u;
    const int ez = (t / 9 + rz) % 3 - 1, ey = ((t / 3) % 3 + ry) % 3 - 1, ex = (t % 3 + rx) % 3 - 1;
    const int64_t s = src[t], p0 = node_p0[s], p1 = node_p0[s + 1];
    for (int64_t k = 0; k < p1 - p0; k++) {
      const int64_t q = descending ? p1 - 1 - k : p0 + k;
      dens = (float)((double)dens + ahf_w9_contrib(&W[q], ex, ey, ez));
    }
  }
  return dens;
}

/* Morton index of an octet coordinate triple (21 bits per axis) */
AHF_HD uint64_t ahf_spread3(uint64_t v)
{
  v &= 0x1FFFFFull;
  v = (v | v << 32) & 0x1F00000000FFFFull;
  v = (v | v << 16) & 0x1F0000FF0000FFull;
  v = (v | v << 8) & 0x100F00F00F00F00Full;
  v = (v | v << 4) & 0x10C30C30C30C30C3ull;
  v = (v | v << 2) & 0x1249249249249249ull;
  return v;
}
AHF_HD uint64_t ahf_compact3(uint64_t v)
{
  v &= 0x1249249249249249ull;
  v = (v ^ (v >> 2)) & 0x10C30C30C30C30C3ull;
  v = (v ^ (v >> 4)) & 0x100F00F00F00F00Full;
  v = (v ^ (v >> 8)) & 0x1F0000FF0000FFull;
  v = (v ^ (v >> 16)) & 0x1F00000000FFFFull;
  v = (v ^ (v >> 32)) & 0x1FFFFFull;
  return v;
}
AHF_HD uint64_t ahf_morton(uint64_t key)
{
  return ahf_spread3((uint64_t)ahf_kx(key)) | ahf_spread3((uint64_t)ahf_ky(key)) << 1 |
         ahf_spread3((uint64_t)ahf_kz(key)) << 2;
}

/* test_node, src/libamr_serial/refine_grid.c:113-138: returns NF_* bits */
AHF_HD uint8_t ahf_node_flags(const level_view *lv, const uint32_t *pmask, const float *dens,
                              double thr, int64_t o, int c)
{
  const int64_t n = 8 * o + c;
  if (pmask[n] != AHF_FULL27) return 0;
  uint8_t f = NF_INTERIOR;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = 0; dx <= 1; dx++) {
        const int64_t m = ahf_nbr_node(lv->nbr, o, c, dx, dy, dz);
        if ((double)dens[m] >= thr) return (uint8_t)(f | NF_TEST);
      }
  return f;
}

/*
 * ahf_presence_mask from values read once per octet: its 27 neighbour
 * octets nb[] (index ahf_e27), and for each (ey,ez) the x-run link bit of
 * the neighbour octets at (0,ey,ez) and (-1,ey,ez) (lk0, lkm, index
 * (ez+1)*3 + (ey+1); read only where those octets exist).  Same rules.
 */
AHF_HD uint32_t ahf_presence_mask_oct(const int32_t *nb, const uint8_t *lk0, const uint8_t *lkm, int64_t ox,
                                      int64_t half, int c)
{
  const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
  uint32_t m = 0;
  for (int dz = -1; dz <= 1; dz++) {
    const int ez = ahf_carry(cz, dz);
    if (nb[ahf_e27(0, 0, ez)] < 0) continue;
    for (int dy = -1; dy <= 1; dy++) {
      const int ey = ahf_carry(cy, dy);
      if (nb[ahf_e27(0, ey, ez)] < 0) continue;
      const int base = (dz + 1) * 9 + (dy + 1) * 3;
      const int li = (ez + 1) * 3 + (ey + 1);
      m |= 1u << (base + 1);
      if (cx == 0 || (nb[ahf_e27(1, ey, ez)] >= 0 && (ox == half - 1 || (lk0[li] & 1)))) m |= 1u << (base + 2);
      if (cx == 1 || (nb[ahf_e27(-1, ey, ez)] >= 0 && (ox == 0 || (lkm[li] & 1)))) m |= 1u << (base + 0);
    }
  }
  return m;
}

/*
 * Block test of ref_pquad / ref_cquad (src/libamr_serial/refine_grid.c:
 * 357-400 and 641-686 choose low/up offsets, :403-455 / :689-751 loop over
 * the block but its last member(s), :491-540 / :777-826 CASE 2 tests one
 * more member at the box edge).  Block = [b0, b0+blen) in level units,
 * idx = member index inside the block.
 */
AHF_HD int ahf_block_tested(int64_t b0, int64_t blen, int64_t idx, int is_last_block, int wrap, int64_t L)
{
  int64_t lo = 1, up = -1;
  if (wrap) {
    if (b0 == 0 && b0 + blen == L) { lo = 0; up = 0; }
    else if (b0 == 0) { lo = 0; up = -1; }
    else if (b0 + blen == L) { lo = 1; up = 0; }
  }
  const int64_t loop_end = blen - 1 + up;
  if (idx >= lo && idx < loop_end) return 1;
  if (is_last_block && b0 + blen == L) {
    const int64_t extra = lo > loop_end ? lo : loop_end;
    if (idx == extra) return 1;
  }
  return 0;
}

/* first/last index of the run of consecutive values around position i of a
 * sorted list (values v(i), contiguous means v(i+1) == v(i)+1) */
#define AHF_RUN_BOUNDS(N, LO_LIMIT, HI_LIMIT, VAL, i, r0, r1)                \
  do {                                                                      \
    r0 = (i);                                                               \
    while (r0 > (LO_LIMIT) && VAL(r0 - 1) == VAL(r0) - 1) r0--;             \
    r1 = (i);                                                               \
    while (r1 + 1 < (HI_LIMIT) && VAL(r1 + 1) == VAL(r1) + 1) r1++;         \
  } while (0)

/* is level-plane Z = 2*plane_z[p] + cz refined by ref_pquad? */
AHF_HD int ahf_plane_tested(const level_view *lv, int64_t p, int cz)
{
  const int64_t half = lv->L / 2;
  int64_t p0, p1;
#define PZ(k) ((int64_t)lv->plane_z[k])
  AHF_RUN_BOUNDS(lv->n_plane, 0, lv->n_plane, PZ, p, p0, p1);
#undef PZ
  const int wrap = lv->plane_z[0] == 0 && lv->plane_z[lv->n_plane - 1] == half - 1;
  const int64_t b0 = 2 * (int64_t)lv->plane_z[p0], blen = 2 * (p1 - p0 + 1);
  const int64_t idx = 2 * (int64_t)lv->plane_z[p] + cz - b0;
  return ahf_block_tested(b0, blen, idx, p1 == lv->n_plane - 1, wrap, lv->L);
}

/* is level-row Y = 2*oy(r)+cy tested by ref_cquad?  The periodic test looks at
 * the first plane of the plane's pquad (refine_grid.c:359-367). */
AHF_HD int ahf_row_tested(const level_view *lv, int64_t r, int cy)
{
  const int64_t half = lv->L / 2;
  const int64_t p = lv->row_plane[r];
  const int64_t q0 = lv->plane_row0[p], q1 = lv->plane_row0[p + 1];
  int64_t r0, r1;
#define RY(k) ahf_row_y(lv->row_key[k])
  AHF_RUN_BOUNDS(q1 - q0, q0, q1, RY, r, r0, r1);
#undef RY
  /* first octet plane of this plane's run */
  int64_t pp = p;
  while (pp > 0 && lv->plane_z[pp - 1] == lv->plane_z[pp] - 1) pp--;
  const int64_t f0 = lv->plane_row0[pp], f1 = lv->plane_row0[pp + 1];
  const int wrap = ahf_row_y(lv->row_key[f0]) == 0 && ahf_row_y(lv->row_key[f1 - 1]) == half - 1;
  const int64_t b0 = 2 * ahf_row_y(lv->row_key[r0]), blen = 2 * (r1 - r0 + 1);
  const int64_t idx = 2 * ahf_row_y(lv->row_key[r]) + cy - b0;
  return ahf_block_tested(b0, blen, idx, r1 == q1 - 1, wrap, lv->L);
}

/*
 * ref_nquad on one level row (src/libamr_serial/refine_grid.c:155-303):
 * walks the row's x-runs and marks the nodes that become octets of the next
 * level (NF_REFINED) and whether the fine run continues to x+1 (NF_LINK).
 * flags[] holds NF_INTERIOR/NF_TEST of every node of the level.
 * Returns the number of refined nodes; *anomaly is set when AHF would emit
 * an empty run (its behaviour there is undefined).
 */
AHF_HD int64_t ahf_refine_row(const level_view *lv, uint8_t *flags, int64_t r, int cy, int cz, int *anomaly)
{
  const int64_t L = lv->L, half = L / 2;
  const int64_t a = lv->row_oct0[r], b = lv->row_oct0[r + 1];
  const int csub = cz * 4 + cy * 2;
  const int last_at_edge = ahf_kx(lv->oct_key[b - 1]) == half - 1;
#define NODE_OF(oi, cx) (8 * (int64_t)(oi) + csub + (cx))
  int64_t nref = 0;
  int64_t i0 = a;
  while (i0 < b) {
    int64_t i1 = i0;
    while (i1 + 1 < b && (lv->oct_link[i1] & 1)) i1++;
    const int64_t X0 = 2 * ahf_kx(lv->oct_key[i0]);
    const int64_t len = 2 * (i1 - i0 + 1);
    const int is_last = (i1 == b - 1);
    int state = 0;
    int64_t start = 0, xlen = 0, xoff = 1;
#define NODE_K(k) NODE_OF(i0 + (k) / 2, (k) & 1)
#define EMIT()                                                                 \
  do {                                                                         \
    if (xlen == 0) *anomaly = 1;                                               \
    for (int64_t u = 0; u < xlen / 2; u++) {                                   \
      const int64_t Xc = start / 2 + u;                                        \
      const int64_t k = Xc - X0 + 0;                                           \
      const int64_t nd = NODE_OF(i0 + k / 2, k & 1);                           \
      flags[nd] |= (uint8_t)(NF_REFINED | (u + 1 < xlen / 2 ? NF_LINK : 0));   \
    }                                                                          \
    nref += xlen / 2;                                                          \
  } while (0)
    if (X0 == 0 && last_at_edge) {
      if (flags[NODE_OF(b - 1, 1)] & NF_TEST) { state = 1; start = 0; }
      xoff = 0;
    }
    int64_t k = xoff;
    for (; k < len - 1; k++) {
      const uint8_t f = flags[NODE_K(k)];
      if (f & NF_TEST) {
        if (!state) { state = 1; start = 2 * (X0 + k); }
        xlen += 2;
      } else if (state) {
        state = 0;
        if (f & NF_INTERIOR) xlen += 2;
        EMIT();
        xlen = 0;
      }
    }
    if (!is_last) {
      if (state) EMIT();
    } else if (X0 + len == L && (flags[NODE_K(k)] & NF_TEST)) {
      if (!state) start = 2 * (X0 + k);
      xlen += 2;
      EMIT();
    } else if (state) {
      EMIT();
    }
#undef EMIT
#undef NODE_K
    i0 = i1 + 1;
  }
#undef NODE_OF
  return nref;
}

/*
 * relink (src/libamr_serial/relink.c:31-288) for one particle of a refined
 * mother node at level coordinates (MX,MY,MZ): the first child, in fine
 * traversal order, that is interior on the fine level and whose closed cell
 * |x - c| <= h holds the particle; -1 if none (the particle stays).
 */
AHF_HD int ahf_relink_child(float px, float py, float pz, int64_t MX, int64_t MY, int64_t MZ,
                            double Lf, uint8_t interior_mask)
{
  const double h = 0.5 / Lf;
  for (int c = 0; c < 8; c++) {
    if (!((interior_mask >> c) & 1u)) continue;
    const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
    const double xc = ((double)(2 * MX + cx) / Lf) + h;
    const double yc = ((double)(2 * MY + cy) / Lf) + h;
    const double zc = ((double)(2 * MZ + cz) / Lf) + h;
    if (fabs((double)px - xc) <= h && fabs((double)py - yc) <= h && fabs((double)pz - zc) <= h) return c;
  }
  return -1;
}

/* ahf_relink_child with the six child-centre coordinates computed once
 * (the same expressions, so the same bits) */
AHF_HD int ahf_relink_child_h(float px, float py, float pz, int64_t MX, int64_t MY, int64_t MZ, double Lf,
                              uint8_t interior_mask)
{
  const double h = 0.5 / Lf;
  const double xc[2] = {((double)(2 * MX) / Lf) + h, ((double)(2 * MX + 1) / Lf) + h};
  const double yc[2] = {((double)(2 * MY) / Lf) + h, ((double)(2 * MY + 1) / Lf) + h};
  const double zc[2] = {((double)(2 * MZ) / Lf) + h, ((double)(2 * MZ + 1) / Lf) + h};
  const int inx[2] = {fabs((double)px - xc[0]) <= h, fabs((double)px - xc[1]) <= h};
  const int iny[2] = {fabs((double)py - yc[0]) <= h, fabs((double)py - yc[1]) <= h};
  const int inz[2] = {fabs((double)pz - zc[0]) <= h, fabs((double)pz - zc[1]) <= h};
  for (int c = 0; c < 8; c++) {
    if (!((interior_mask >> c) & 1u)) continue;
    if (inx[c & 1] && iny[(c >> 1) & 1] && inz[c >> 2]) return c;
  }
  return -1;
}

/* position of level-row (r,cy,cz) in (Z,Y) raster order of the level's rows */
AHF_HD int64_t ahf_row_rank(const level_view *lv, int64_t r, int cy, int cz)
{
  const int64_t p = lv->row_plane[r];
  const int64_t q0 = lv->plane_row0[p], nr = lv->plane_row0[p + 1] - q0;
  return 4 * q0 + cz * 2 * nr + 2 * (r - q0) + cy;
}

/*
 * Exact parallel form of AHF's density accumulation d = (float)((double)d + c)
 * (src/libamr_serial/density.c:362-383) over a run of contributions.
 *
 * While |d| stays in one binade [2^e, 2^(e+1)), d = K*u with u = 2^(e-23)
 * and K an integer, 2^23 <= |K| < 2^24.  If the double sum d + c also lies
 * in that binade, its rounding step is v = u*2^-29, and since K*u is an
 * even multiple of v, fl64(K*u + c) = K*u + v*R with R = rint(c/v), which
 * depends on c alone.  The float rounding then gives K' = K + r, where r is
 * R*2^-29 rounded to nearest, and on an exact tie the r that makes K' even.
 * So a run of steps inside one binade is a prefix sum of integers; only a
 * tie needs the parity of the running K.  A step that leaves the binade (or
 * starts from zero or a subnormal) is taken as the plain FP32 step.
 *
 * ahf_fold_exponent: e of a normal float (0 for zero, subnormal, inf, NaN,
 * with *ok = 0).  ahf_fold_lane: for one contribution c in binade e, R as an
 * integer (ok = 0 when c is not finite or R would not fit), the rounded
 * increment r for the non-tie case, and whether it is a tie (then r is the
 * lower candidate).
 */
AHF_HD int ahf_fold_exponent(float d, int *ok)
{
  union { float f; uint32_t u; } b;
  b.f = d;
  const int ex = (int)((b.u >> 23) & 0xFFu);
  *ok = ex != 0 && ex != 0xFF;
  return ex - 127;
}

AHF_HD int ahf_fold_lane(double c, int e, int64_t *R, int64_t *r, int *tie)
{
  *R = 0;
  *r = 0;
  *tie = 0;
  if (!(c == c) || c > 1e300 || c < -1e300) return 0;
  const double q = rint(ldexp(c, 52 - e)); /* c / v, exact scaling, then rounded to nearest even */
  if (q >= 4611686018427387904.0 || q <= -4611686018427387904.0) return 0; /* |R| < 2^62 */
  const int64_t Ri = (int64_t)q;
  const int64_t fl = Ri >= 0 ? Ri / ((int64_t)1 << 29) : -((-Ri + ((int64_t)1 << 29) - 1) / ((int64_t)1 << 29));
  const int64_t rem = Ri - fl * ((int64_t)1 << 29); /* floor division: rem in [0, 2^29) */
  *R = Ri;
  *r = fl + (rem > ((int64_t)1 << 28) ? 1 : 0);
  *tie = rem == ((int64_t)1 << 28);
  return 1;
}

/* K (the integer of d in its binade) and the binade tests of one step */
AHF_HD int64_t ahf_fold_K(float d, int e) { return (int64_t)ldexp((double)d, 23 - e); }
AHF_HD int ahf_fold_K_in(int64_t K, int neg)
{
  return neg ? (K > -((int64_t)1 << 24) && K <= -((int64_t)1 << 23)) : (K >= ((int64_t)1 << 23) && K < ((int64_t)1 << 24));
}
/* the double sum T*v = K*u + c stays inside the binade with a margin of one
 * rounding step, so its rounding step is v */
AHF_HD int ahf_fold_T_in(int64_t K, int64_t R, int neg)
{
  const int64_t T = K * ((int64_t)1 << 29) + R;
  const int64_t lo = ((int64_t)1 << 52) + 1, hi = ((int64_t)1 << 53) - 1;
  return neg ? (T <= -lo && T >= -hi) : (T >= lo && T <= hi);
}

#endif /* AHF_AMR_CORE_H */
