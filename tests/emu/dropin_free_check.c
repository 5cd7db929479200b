/*
 * dropin_free_check.c -- test-only: the drop-in's grids released by AHF's
 * own free_grid (src/libutility/alloc_struct.c:584-605), on the host under
 * AddressSanitizer.
 *
 * Linked from ahf_amd/csrc/dropin.c (the drop-in's host code as the library
 * builds it, with AddressSanitizer), the host stand-in of the device API
 * (emu_device.c: the replay of the device pipeline) and the reference's own
 * objects (AHF's globals and free_grid; the functions the drop-in replaces
 * made local, as in the integration recipe).  The program runs
 * gen_domgrids, ll, zero_dens and gen_AMRhierarchy through the drop-in,
 * checks that every grid holds the nodes the build reports and that every
 * particle is listed exactly once over all grids, then frees the grids from
 * the finest down with free_grid as src/libahf/ahf_halos.c:470-484 does.
 * AddressSanitizer fails the run on any free of memory the drop-in did not
 * allocate the way AHF's c_pquad/c_cquad/c_nquad/c_node do, and its leak
 * check on any block that free_grid leaves behind.
 *
 *   dropin_free_check <snapshot>     prints {"grids": G, "nodes": N}
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../ahf_amd/csrc/snapfile.h"
#include "../../include/ahf_dropin.h"

/* AHF's own (src/libutility/alloc_struct.h:30) */
void free_grid(gridls *cur_grid, int *no_grids);

static void die(const char *m)
{
  fprintf(stderr, "dropin_free_check: %s\n", m);
  exit(2);
}

int main(int argc, char **argv)
{
  if (argc < 2) die("usage: dropin_free_check <snapshot>");
  snap_data s;
  if (snap_read(argv[1], &s)) die("cannot read the snapshot");
  const uint64_t n = s.h.npart;
  /* the parameters the path reads (src/startrun.c:499-547) */
  memset(&simu, 0, sizeof(simu));
  simu.NGRID_DOM = (int)s.h.L_dom;
  simu.NGRID_MIN = (int)s.h.L_dom;
  simu.NGRID_MAX = (int)s.h.L_max;
  simu.Nth_dom = s.h.Nth_dom;
  simu.Nth_ref = s.h.Nth_ref;
  simu.no_part = n;
  simu.no_vpart = (double)n;
  simu.mean_dens = 1.0;
  memset(&global, 0, sizeof(global));
  global.fst_cycle = 1;
  partptr P = (partptr)calloc(n ? n : 1, sizeof(part));
  if (!P) die("out of memory");
  for (uint64_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      P[i].pos[d] = s.pos[3 * i + d];
      P[i].mom[d] = s.mom[3 * i + d];
    }
  int no_grids;
  gridls *grid_list = gen_domgrids(&no_grids);
  ll(n, P, global.dom_grid);
  global.fst_part = P;
  global.no_part = n;
  zero_dens(global.dom_grid);
  global.fst_cycle = 1;
  gen_AMRhierarchy(&grid_list, &no_grids);

  /* every grid as the build reports it; every particle listed once */
  uint8_t *seen = (uint8_t *)calloc(n ? n : 1, 1);
  if (!seen) die("out of memory");
  uint64_t nodes = 0, listed = 0;
  for (int g = 0; g < no_grids; g++) {
    uint64_t ng = 0;
    for (pqptr pq = grid_list[g].pquad; pq; pq = pq->next)
      for (cqptr cq0 = pq->loc; cq0 < pq->loc + pq->length; cq0++)
        for (cqptr cq = cq0; cq; cq = cq->next)
          for (nqptr nq0 = cq->loc; nq0 < cq->loc + cq->length; nq0++)
            for (nqptr nq = nq0; nq; nq = nq->next)
              for (nptr nd = nq->loc; nd < nq->loc + nq->length; nd++) {
                ng++;
                for (partptr p = nd->ll; p; p = p->ll) {
                  const uint64_t i = (uint64_t)(p - P);
                  if (i >= n || seen[i]) die("a particle is listed twice or out of range");
                  seen[i] = 1;
                  listed++;
                }
              }
    if (ng != grid_list[g].size.no_nodes) die("a grid's node count differs from the build's");
    nodes += ng;
  }
  if (listed != n) die("not every particle is listed");
  free(seen);

  /* src/libahf/ahf_halos.c:470-484: the finest grid first */
  const int total = no_grids;
  int left = no_grids - 1;
  for (gridls *g = grid_list + left; left >= 0; g--) free_grid(g, &left);
  free(grid_list);
  free(P);
  ahf_gpu_finalize();
  snap_free(&s);
  printf("{\"grids\": %d, \"nodes\": %llu}\n", total, (unsigned long long)nodes);
  return 0;
}
