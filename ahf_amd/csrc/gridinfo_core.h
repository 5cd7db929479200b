/*
 * gridinfo_core.h -- per-octet rules of the ahf_gridinfo pass
 * (src/libahf/ahf_gridinfo.c:59-860), shared by the HIP kernels
 * (gridinfo_gpu.hip) and their host replay; plain C99 with AHF_HD.
 *
 * AHF visits a grid's nodes in traversal order (planes of ascending z, rows
 * of ascending y, ascending x) and joins each node to its six face
 * neighbours as get_TSCnodes finds them (colourInfo, :1126-1310).  With the
 * presence rules of ahf_presence_mask, on a grid of octets:
 *   - the eight nodes of an octet are face neighbours of each other (the
 *     node rows and columns inside an octet always exist), so a connected
 *     component is a union of whole octets;
 *   - between octets, the y and z face neighbours exist exactly where the
 *     neighbouring octet does (periodic wraps included), and the x face
 *     neighbours where the x-run links the two octets or, across the
 *     periodic wrap, where the octet on the far side exists;
 *   - a node other than the octet's first (cx = cy = cz = 0) has a face
 *     neighbour before it in its own octet, and the first node's earlier
 *     neighbours are those of the three octets before it in x (linked, not
 *     across the wrap), y and z (not across the wraps), whose octet indices
 *     are lower: colours start only at first nodes, in octet order.
 */
#ifndef AHF_GRIDINFO_CORE_H
#define AHF_GRIDINFO_CORE_H

#include "amr_core.h"

/* the octets joined to octet o in -x, -y and -z (-1 for none): through
 * the run link, or the wraps at ox = 0, oy = 0, oz = 0.  Every edge of the
 * octet graph is one of these for exactly one of its two octets. */
AHF_HD void ahf_gi_oct_edges(const level_view *v, int64_t o, int64_t e[3])
{
  const int32_t *nb = v->nbr + 27 * o;
  const int32_t xm = nb[ahf_e27(-1, 0, 0)];
  e[0] = xm >= 0 && (ahf_kx(v->oct_key[o]) == 0 || (v->oct_link[xm] & 1u)) ? (int64_t)xm : -1;
  e[1] = nb[ahf_e27(0, -1, 0)];
  e[2] = nb[ahf_e27(0, 0, -1)];
}

/* the octet's first node starts a colour: no joined octet before it in x,
 * y or z other than across a wrap */
AHF_HD int ahf_gi_oct_start(const level_view *v, int64_t o, const int64_t e[3])
{
  const uint64_t k = v->oct_key[o];
  return !((e[0] >= 0 && ahf_kx(k) > 0) || (e[1] >= 0 && ahf_ky(k) > 0) || (e[2] >= 0 && ahf_kz(k) > 0));
}

/* testBound (ahf_gridinfo.c:1090-1124, used at :604-690): bit 0/1/2 when
 * the octet's nodes on the low x/y/z face of the box find the neighbour
 * across it */
AHF_HD uint32_t ahf_gi_oct_periodic(const level_view *v, int64_t o, const int64_t e[3])
{
  const uint64_t k = v->oct_key[o];
  return (uint32_t)(ahf_kx(k) == 0 && e[0] >= 0) | (uint32_t)(ahf_ky(k) == 0 && e[1] >= 0) << 1 |
         (uint32_t)(ahf_kz(k) == 0 && e[2] >= 0) << 2;
}

/* the x-run segment starts at o: no link from the octet before it */
AHF_HD int ahf_gi_oct_seg_start(const level_view *v, int64_t o, const int64_t e[3])
{
  return !(e[0] >= 0 && ahf_kx(v->oct_key[o]) > 0);
}

#endif
