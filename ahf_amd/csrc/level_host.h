/*
 * level_host.h -- host-side copy of one AMR level as produced by the device
 * pipeline, and the AHFGRID1 dump of a hierarchy (the layout written by
 * oracle/ref_driver.c, so GPU results compare field by field with the
 * reference and with the CPU oracle).
 */
#ifndef AHF_LEVEL_HOST_H
#define AHF_LEVEL_HOST_H

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "octet_raster.h"

typedef struct {
  int64_t L;
  int64_t n_oct;
  double critdens, masstopartdens;
  uint64_t size_no_part, size_no_nodes;
  uint64_t *oct_key;   /* n_oct */
  uint8_t *oct_link;   /* n_oct */
  float *dens;         /* 8*n_oct, octet-major */
  int64_t *list0;      /* 8*n_oct+1: final per-node lists (CSR) */
  uint32_t *list;      /* particle indices in AHF list order */
} host_level;

static inline void host_level_free(host_level *h)
{
  free(h->oct_key); free(h->oct_link); free(h->dens); free(h->list0); free(h->list);
  memset(h, 0, sizeof(*h));
}

static inline int ahf_put(FILE *f, const void *p, size_t sz, size_t n)
{
  return (n == 0 || fwrite(p, sz, n, f) == n) ? 0 : -1;
}

/* write one level in AHFGRID1 layout */
static inline int host_level_dump(FILE *f, const host_level *h)
{
  octet_raster_t r;
  if (octet_raster_build(h->oct_key, h->oct_link, h->n_oct, &r)) return -1;
  int bad = 0;
  uint64_t n_listed = (uint64_t)h->list0[8 * h->n_oct];
  uint64_t hdr[3] = {(uint64_t)h->L, h->size_no_nodes, h->size_no_part};
  double dh[2] = {h->critdens, h->masstopartdens};
  uint64_t cnt[5] = {(uint64_t)r.n_nq, (uint64_t)r.n_node, n_listed, (uint64_t)r.n_pq, (uint64_t)r.n_cq};
  bad |= ahf_put(f, hdr, 8, 3);
  bad |= ahf_put(f, dh, 8, 2);
  bad |= ahf_put(f, cnt, 8, 5);
  bad |= ahf_put(f, r.nq, 4, (size_t)(4 * r.n_nq));
  bad |= ahf_put(f, r.pq, 4, (size_t)(2 * r.n_pq));
  bad |= ahf_put(f, r.cq, 4, (size_t)(3 * r.n_cq));
  for (int64_t i = 0; i < r.n_node && !bad; i++) bad |= ahf_put(f, &h->dens[r.order[i]], 4, 1);
  for (int64_t i = 0; i < r.n_node && !bad; i++) {
    uint32_t c = (uint32_t)(h->list0[r.order[i] + 1] - h->list0[r.order[i]]);
    bad |= ahf_put(f, &c, 4, 1);
  }
  for (int64_t i = 0; i < r.n_node && !bad; i++)
    for (int64_t q = h->list0[r.order[i]]; q < h->list0[r.order[i] + 1]; q++) {
      uint64_t v = h->list[q];
      bad |= ahf_put(f, &v, 8, 1);
    }
  octet_raster_free(&r);
  return bad ? -1 : 0;
}

static inline int hierarchy_dump(const char *path, const host_level *lv, int nlev)
{
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  int bad = ahf_put(f, "AHFGRID1", 1, 8);
  int32_t ng[2] = {nlev, 0};
  bad |= ahf_put(f, ng, 4, 2);
  for (int i = 0; i < nlev && !bad; i++) bad |= host_level_dump(f, &lv[i]);
  bad |= fclose(f) != 0;
  return bad ? -1 : 0;
}

#endif
