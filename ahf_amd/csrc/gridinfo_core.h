/*
 * gridinfo_core.h -- per-node rules of the ahf_gridinfo pass
 * (src/libahf/ahf_gridinfo.c:59-860), shared by the HIP kernels
 * (gridinfo_gpu.hip) and their host replay; plain C99 with AHF_HD.
 *
 * AHF visits a grid's nodes in traversal order: planes of ascending z, rows
 * of ascending y within a plane, ascending x within a row.  A node's face
 * neighbours are the six that get_TSCnodes finds (the presence bits of
 * ahf_presence_mask), in colourInfo's order -y, -x, -z, +y, +x, +z
 * (ahf_gridinfo.c:1126-1310).
 */
#ifndef AHF_GRIDINFO_CORE_H
#define AHF_GRIDINFO_CORE_H

#include "amr_core.h"

/* presence bits (dz+1)*9 + (dy+1)*3 + (dx+1) of the six face neighbours */
#define AHF_GI_MX 12
#define AHF_GI_PX 14
#define AHF_GI_MY 10
#define AHF_GI_PY 16
#define AHF_GI_MZ 4
#define AHF_GI_PZ 22

/* traversal rank of node n = 8*o + c: the nodes of the octet planes before
 * (8 each), of the node plane cz = 0 of this octet plane when cz = 1 (4 per
 * octet of the plane), of the octet rows before in the plane (4 each), of
 * the node row cy = 0 of this octet row when cy = 1 (2 per octet of the
 * row), of the octets before in the row (2 each), and cx */
AHF_HD int64_t ahf_gi_row_of(const level_view *v, int64_t o)
{
  int64_t lo = 0, hi = v->n_row - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (v->row_oct0[mid] <= o) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

/* rank of node c of octet o in octet row r */
AHF_HD int64_t ahf_gi_rank_in(const level_view *v, int64_t o, int64_t r, int c)
{
  const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
  const int64_t p = v->row_plane[r];
  const int64_t pb = v->row_oct0[v->plane_row0[p]], pn = v->row_oct0[v->plane_row0[p + 1]] - pb;
  const int64_t rb = v->row_oct0[r], rn = v->row_oct0[r + 1] - rb;
  return 8 * pb + cz * 4 * pn + 4 * (rb - pb) + cy * 2 * rn + 2 * (o - rb) + cx;
}

AHF_HD int64_t ahf_gi_rank(const level_view *v, int64_t n)
{
  return ahf_gi_rank_in(v, n >> 3, ahf_gi_row_of(v, n >> 3), (int)(n & 7));
}

/* the found face neighbours of node n (node index, or -1 where
 * get_TSCnodes finds none), in the order -x, +x, -y, +y, -z, +z */
AHF_HD void ahf_gi_faces(const level_view *v, uint32_t m, int64_t n, int64_t u[6])
{
  const int64_t o = n >> 3;
  const int c = (int)(n & 7);
  u[0] = (m >> AHF_GI_MX) & 1u ? ahf_nbr_node(v->nbr, o, c, -1, 0, 0) : -1;
  u[1] = (m >> AHF_GI_PX) & 1u ? ahf_nbr_node(v->nbr, o, c, 1, 0, 0) : -1;
  u[2] = (m >> AHF_GI_MY) & 1u ? ahf_nbr_node(v->nbr, o, c, 0, -1, 0) : -1;
  u[3] = (m >> AHF_GI_PY) & 1u ? ahf_nbr_node(v->nbr, o, c, 0, 1, 0) : -1;
  u[4] = (m >> AHF_GI_MZ) & 1u ? ahf_nbr_node(v->nbr, o, c, 0, 0, -1) : -1;
  u[5] = (m >> AHF_GI_PZ) & 1u ? ahf_nbr_node(v->nbr, o, c, 0, 0, 1) : -1;
}

/* testBound (ahf_gridinfo.c:1090-1124, used at :604-690): bit 0/1/2 when
 * the node lies on the low x/y/z face of the box and the neighbour across
 * that face is found */
AHF_HD uint32_t ahf_gi_periodic(const level_view *v, uint32_t m, int64_t n)
{
  const int64_t o = n >> 3;
  const int c = (int)(n & 7);
  const uint64_t k = v->oct_key[o];
  const int64_t x = 2 * ahf_kx(k) + (c & 1), y = 2 * ahf_ky(k) + ((c >> 1) & 1), z = 2 * ahf_kz(k) + (c >> 2);
  uint32_t f = 0;
  if (x == 0 && ((m >> AHF_GI_MX) & 1u)) f |= 1u;
  if (y == 0 && ((m >> AHF_GI_MY) & 1u)) f |= 2u;
  if (z == 0 && ((m >> AHF_GI_MZ) & 1u)) f |= 4u;
  return f;
}

#endif
