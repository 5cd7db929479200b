/*
 * octet_raster.h -- AHF's traversal order of a level held as octets.
 *
 * Given the octets of a level (sorted by oz, oy, ox; see amr_core.h) and
 * their x-run links, produce what AHF's quads would hold
 * (src/tdef.h:144-183, traversal as in src/libamr_serial/density.c:276-313):
 *   - pquads: runs of consecutive present planes (z0, length);
 *   - cquads: per plane, runs of consecutive present rows (z, y0, length);
 *   - nquads: per row, the linked runs of nodes (z, y, x0, length);
 *   - order:  octet-major node index of every node in traversal order.
 * Host code: used to rebuild AHF's gridls from a device level and to dump it.
 */
#ifndef AHF_OCTET_RASTER_H
#define AHF_OCTET_RASTER_H

#include <stdint.h>
#include <stdlib.h>

typedef struct {
  int64_t n_pq, n_cq, n_nq, n_node;
  int32_t *pq;   /* n_pq x 2 */
  int32_t *cq;   /* n_cq x 3 */
  int32_t *nq;   /* n_nq x 4 */
  int64_t *order;/* n_node */
} octet_raster_t;

static inline void octet_raster_free(octet_raster_t *r)
{
  free(r->pq); free(r->cq); free(r->nq); free(r->order);
  r->pq = r->cq = r->nq = 0;
  r->order = 0;
  r->n_pq = r->n_cq = r->n_nq = r->n_node = 0;
}

#define OR_KX(k) ((int64_t)((k) & 0x1FFFFFull))
#define OR_KY(k) ((int64_t)(((k) >> 21) & 0x1FFFFFull))
#define OR_KZ(k) ((int64_t)(((k) >> 42) & 0x1FFFFFull))

/* returns 0, or -1 on allocation failure */
static inline int octet_raster_build(const uint64_t *key, const uint8_t *link, int64_t n_oct,
                                     octet_raster_t *r)
{
  /* octet rows [row0[i], row0[i+1]) and planes [pl0[j], pl0[j+1]) of rows */
  int64_t *row0 = (int64_t *)malloc((size_t)(n_oct + 1) * sizeof(int64_t));
  int64_t *pl0 = (int64_t *)malloc((size_t)(n_oct + 1) * sizeof(int64_t));
  int64_t nrow = 0, npl = 0;
  r->pq = r->cq = r->nq = 0;
  r->order = 0;
  if (!row0 || !pl0) { free(row0); free(pl0); return -1; }
  for (int64_t i = 0; i < n_oct; i++) {
    if (i == 0 || OR_KY(key[i]) != OR_KY(key[i - 1]) || OR_KZ(key[i]) != OR_KZ(key[i - 1])) {
      if (i == 0 || OR_KZ(key[i]) != OR_KZ(key[i - 1])) pl0[npl++] = nrow;
      row0[nrow++] = i;
    }
  }
  row0[nrow] = n_oct;
  pl0[npl] = nrow;
  /* upper bounds for the outputs */
  int64_t n_runs = 0;
  for (int64_t i = 0; i < n_oct; i++) {
    int run_end = (i + 1 == n_oct) || !(link[i] & 1) ||
                  OR_KY(key[i + 1]) != OR_KY(key[i]) || OR_KZ(key[i + 1]) != OR_KZ(key[i]);
    n_runs += run_end;
  }
  r->pq = (int32_t *)malloc((size_t)(npl + 1) * 2 * sizeof(int32_t));
  r->cq = (int32_t *)malloc((size_t)(2 * nrow + 1) * 3 * sizeof(int32_t));
  r->nq = (int32_t *)malloc((size_t)(4 * n_runs + 1) * 4 * sizeof(int32_t));
  r->order = (int64_t *)malloc((size_t)(8 * n_oct + 1) * sizeof(int64_t));
  if (!r->pq || !r->cq || !r->nq || !r->order) {
    free(row0); free(pl0); octet_raster_free(r);
    return -1;
  }
  int64_t npq = 0, ncq = 0, nnq = 0, nnode = 0;
  for (int64_t j = 0; j < npl; j++) {
    int64_t oz = OR_KZ(key[row0[pl0[j]]]);
    if (j == 0 || OR_KZ(key[row0[pl0[j - 1]]]) != oz - 1) {
      r->pq[2 * npq] = (int32_t)(2 * oz);
      r->pq[2 * npq + 1] = 0;
      npq++;
    }
    r->pq[2 * (npq - 1) + 1] += 2;
    for (int cz = 0; cz < 2; cz++) {
      int64_t Z = 2 * oz + cz;
      for (int64_t i = pl0[j]; i < pl0[j + 1]; i++) {
        int64_t oy = OR_KY(key[row0[i]]);
        if (i == pl0[j] || OR_KY(key[row0[i - 1]]) != oy - 1) {
          r->cq[3 * ncq] = (int32_t)Z;
          r->cq[3 * ncq + 1] = (int32_t)(2 * oy);
          r->cq[3 * ncq + 2] = 0;
          ncq++;
        }
        r->cq[3 * (ncq - 1) + 2] += 2;
        for (int cy = 0; cy < 2; cy++) {
          int64_t Y = 2 * oy + cy;
          int64_t a = row0[i], b = row0[i + 1];
          int64_t s = a;
          while (s < b) {
            int64_t e = s;
            while (e + 1 < b && (link[e] & 1)) e++;
            r->nq[4 * nnq] = (int32_t)Z;
            r->nq[4 * nnq + 1] = (int32_t)Y;
            r->nq[4 * nnq + 2] = (int32_t)(2 * OR_KX(key[s]));
            r->nq[4 * nnq + 3] = (int32_t)(2 * (e - s + 1));
            nnq++;
            for (int64_t o = s; o <= e; o++)
              for (int cx = 0; cx < 2; cx++) r->order[nnode++] = 8 * o + cz * 4 + cy * 2 + cx;
            s = e + 1;
          }
        }
      }
    }
  }
  r->n_pq = npq;
  r->n_cq = ncq;
  r->n_nq = nnq;
  r->n_node = nnode;
  free(row0);
  free(pl0);
  return 0;
}

#endif
