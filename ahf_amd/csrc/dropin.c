/*
 * dropin.c -- AHF's hot-path entry points (include/ahf_dropin.h) with AHF's
 * prototypes, structs and globals, implemented on the device API of
 * include/ahf_gpu.h.  Host C: this file only converts between AHF's
 * pointer structures (linked quads, AoS particles, linked particle lists,
 * HALO member arrays) and the library's flat device arrays; every deposit,
 * refinement, relink, sort and unbinding runs on the GPU.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "ahf_dropin.h"
#include "ahf_gpu.h"
#include "octet_raster.h"

/* AHF defines these (src/common.c:23-34, src/libahf/ahf_halos.c:163); the
 * weak copies only let the library link on its own */
__attribute__((weak)) param_simu simu;
__attribute__((weak)) info_global global;
__attribute__((weak)) global_info_t global_info;
__attribute__((weak)) double r_fac, v_fac, phi_fac, rho_fac, Hubble, m_fac, x_fac;

static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static ahfgpu_ctx *g_ctx;
static int g_device;
static const part *g_res_parts; /* particle set resident on the device */
static uint64_t g_res_n;
static uint64_t g_upload;        /* uploads so far; keys belong to one */
static uint64_t g_keys_upload;   /* upload whose keys are resident, 0: none */

static void fatal(const char *what)
{
  fprintf(stderr, "ahf_amd: %s: %s\n", what, g_ctx ? ahfgpu_last_error(g_ctx) : "no device context");
  fflush(stderr);
  exit(1);
}

static void *xcalloc(size_t n, size_t sz)
{
  void *p = calloc(n ? n : 1, sz);
  if (!p) fatal("out of host memory");
  return p;
}

/* AHF's functions read no stage timings: the drop-in's context runs
 * without them */
static ahfgpu_ctx *ctx_get(void)
{
  if (!g_ctx) {
    if (ahfgpu_create(&g_ctx, g_device)) {
      g_ctx = NULL;
      fatal("cannot create a GPU context");
    }
    ahfgpu_set_timing(g_ctx, 0);
  }
  return g_ctx;
}

int ahf_gpu_init(int device)
{
  pthread_mutex_lock(&g_lock);
  int rc = 0;
  if (g_ctx && g_device != device) {
    ahfgpu_destroy(g_ctx);
    g_ctx = NULL;
  }
  g_device = device;
  g_res_parts = NULL;
  g_res_n = 0;
  if (!g_ctx) rc = ahfgpu_create(&g_ctx, device);
  if (rc)
    g_ctx = NULL;
  else
    ahfgpu_set_timing(g_ctx, 0);
  pthread_mutex_unlock(&g_lock);
  return rc;
}

void ahf_gpu_finalize(void)
{
  pthread_mutex_lock(&g_lock);
  if (g_ctx) ahfgpu_destroy(g_ctx);
  g_ctx = NULL;
  g_res_parts = NULL;
  g_res_n = 0;
  pthread_mutex_unlock(&g_lock);
}

/* copy global.fst_part (AoS) to the device as SoA, unless already there */
static void ensure_particles(int force)
{
  const uint64_t n = global.no_part;
  if (!force && g_res_parts == global.fst_part && g_res_n == n) return;
  if (!global.fst_part || n == 0) fatal("no particles (global.fst_part)");
  float *pos = (float *)malloc(n * 3 * sizeof(float)), *mom = (float *)malloc(n * 3 * sizeof(float));
  if (!pos || !mom) fatal("out of host memory");
  const part *P = global.fst_part;
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      pos[3 * i + d] = P[i].pos[d];
      mom[3 * i + d] = P[i].mom[d];
    }
  int rc = ahfgpu_set_particles(ctx_get(), pos, mom, n);
  free(pos);
  free(mom);
  if (rc) fatal("particle upload");
  g_res_parts = global.fst_part;
  g_res_n = n;
  g_upload++;
}

/* AHF's SFC keys of the resident particles (part.sfckey, set and sorted by
 * src/main.c:341-356), for the halo gather */
static void ensure_keys(void)
{
  if (global_info.fst_part != global.fst_part || global_info.no_part != global.no_part)
    fatal("global_info and global hold different particle arrays");
  ensure_particles(0);
  if (g_keys_upload == g_upload) return;
  const uint64_t n = global_info.no_part;
  uint64_t *k = (uint64_t *)malloc(n * sizeof(uint64_t));
  if (!k) fatal("out of host memory");
  for (uint64_t i = 0; i < n; i++) k[i] = global_info.fst_part[i].sfckey;
  int rc = ahfgpu_set_sfc_keys(ctx_get(), k, n);
  free(k);
  if (rc) fatal("SFC keys");
  g_keys_upload = g_upload;
}

/* ------------------------------------------------------------ grids ---- */

/* alloc_quads() layout of a full periodic grid (src/libutility/alloc_struct.c:476-510) */
static void full_quads(gridls *g, long l1dim)
{
  pqptr pq = (pqptr)xcalloc(1, sizeof(pquad));
  g->pquad = pq;
  pq->length = (int)l1dim;
  pq->loc = (cqptr)xcalloc((size_t)l1dim, sizeof(cquad));
  for (long z = 0; z < l1dim; z++) {
    cqptr cq = pq->loc + z;
    cq->length = (int)l1dim;
    cq->loc = (nqptr)xcalloc((size_t)l1dim, sizeof(nquad));
    for (long y = 0; y < l1dim; y++) {
      cq->loc[y].length = (int)l1dim;
      cq->loc[y].loc = (nptr)xcalloc((size_t)l1dim, sizeof(node));
    }
  }
  g->no_pquad = 1;
  g->pquad_array = (pqptr *)xcalloc(1, sizeof(pqptr));
  g->pquad_array[0] = g->pquad;
}

gridls *gen_domgrids(int *no_grids)
{
  double frac = (double)simu.NGRID_DOM / (double)simu.NGRID_MIN;
  int e = (int)(float)(log(frac) / log(2.));
  *no_grids = e + 1;
  gridls *list = (gridls *)xcalloc((size_t)*no_grids, sizeof(gridls));
  gridls *g = list;
  for (long l1 = simu.NGRID_MIN; l1 <= simu.NGRID_DOM; l1 *= 2, g++) {
    double dl = (double)l1;
    g->timecounter = global.super_t;
    g->l1dim = (long unsigned)l1;
    g->spacing = 1.0 / dl;
    g->spacing2 = g->spacing * g->spacing;
    g->masstodens = (dl * dl * dl) / simu.no_vpart;
    g->masstopartdens = (dl * dl * dl) / (double)simu.no_part;
    g->critdens = simu.Nth_dom * g->masstodens;
    full_quads(g, l1);
  }
  global.dom_grid = list + (*no_grids - 1);
  global.domgrid_no = *no_grids - 1;
  return list;
}

/* AHF's head insertion in particle order (src/libamr_serial/lltools.c:32):
 * every node's list runs from its last particle to its first, then what the
 * node held before.  The node of every particle is found in parallel; then
 * each thread owns a range of planes and inserts, in particle order, the
 * particles of its planes, so every list comes out as the sequential loop
 * leaves it. */
void ll(long unsigned npart, partptr fst_part, gridls *cur_grid)
{
  const long unsigned L = cur_grid->l1dim;
  pqptr pq = cur_grid->pquad;
  uint32_t *zs = (uint32_t *)malloc((npart ? npart : 1) * sizeof(uint32_t));
  uint64_t *yx = (uint64_t *)malloc((npart ? npart : 1) * sizeof(uint64_t));
  if (!zs || !yx) fatal("out of host memory");
#pragma omp parallel
  {
#pragma omp for schedule(static)
    for (long unsigned j = 0; j < npart; j++) {
      const partptr p = fst_part + j;
      unsigned long n[3];
      for (int d = 0; d < 3; d++) {
        n[d] = (unsigned long)((double)L * p->pos[d]);
        if (n[d] > L - 1) n[d] = 0;
      }
      zs[j] = (uint32_t)n[2];
      yx[j] = (uint64_t)n[1] * L + n[0];
    }
#ifdef _OPENMP
    const long unsigned T = (long unsigned)omp_get_num_threads(), t = (long unsigned)omp_get_thread_num();
#else
    const long unsigned T = 1, t = 0;
#endif
    const uint32_t z0 = (uint32_t)(L * t / T), z1 = (uint32_t)(L * (t + 1) / T);
    for (long unsigned j = 0; z0 < z1 && j < npart; j++) {
      const uint32_t z = zs[j];
      if (z < z0 || z >= z1) continue;
      const partptr p = fst_part + j;
      nptr nd = pq->loc[z].loc[yx[j] / L].loc + yx[j] % L;
      p->ll = nd->ll;
      nd->ll = p;
    }
  }
  free(zs);
  free(yx);
}

#define FOR_NODES(g, BODY)                                                   \
  for (pqptr pq_ = (g)->pquad; pq_; pq_ = pq_->next)                          \
    for (cqptr cq0_ = pq_->loc; cq0_ < pq_->loc + pq_->length; cq0_++)        \
      for (cqptr cq_ = cq0_; cq_; cq_ = cq_->next)                            \
        for (nqptr nq0_ = cq_->loc; nq0_ < cq_->loc + cq_->length; nq0_++)    \
          for (nqptr nq_ = nq0_; nq_; nq_ = nq_->next)                        \
            for (nptr nd = nq_->loc; nd < nq_->loc + nq_->length; nd++) { BODY }

void zero_dens(gridls *cur_grid)
{
  FOR_NODES(cur_grid, nd->dens = (flouble)(-simu.mean_dens); nd->force.tempdens = 0.0f;)
}

/* a grid read into flat arrays: octets (sorted), links, the node of every
 * octet-major index, and the per-node lists in list order */
typedef struct {
  int64_t n_oct, n_node, n_listed;
  uint64_t *key;
  uint8_t *link;
  nptr *node_of;   /* octet-major index -> node */
  int64_t *list0;
  uint32_t *list;
} flat_grid;

static void flat_free(flat_grid *f)
{
  free(f->key); free(f->link); free(f->node_of); free(f->list0); free(f->list);
  memset(f, 0, sizeof(*f));
}

static int cmp_u64(const void *a, const void *b)
{
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : x > y;
}

static int64_t find_key(const uint64_t *k, int64_t n, uint64_t want)
{
  int64_t lo = 0, hi = n - 1;
  while (lo <= hi) {
    int64_t m = (lo + hi) / 2;
    if (k[m] == want) return m;
    if (k[m] < want) lo = m + 1; else hi = m - 1;
  }
  return -1;
}

#define PK(x, y, z) ((uint64_t)(x) | ((uint64_t)(y) << 21) | ((uint64_t)(z) << 42))

/* every AHF grid is a union of 2x2x2 node blocks (runs of even start and
 * length, mirrored rows and planes); returns 0, or -1 if this one is not.
 * With lists == 0 the particle lists are not read (list0 stays zero and
 * list NULL). */
static int flatten_grid(gridls *g, flat_grid *f, int lists)
{
  memset(f, 0, sizeof(*f));
  int64_t n = 0;
  FOR_NODES(g, (void)nd; n++;)
  f->n_node = n;
  if (n == 0 || n % 8) return -1;
  f->n_oct = n / 8;
  f->key = (uint64_t *)xcalloc((size_t)f->n_oct, 8);
  f->link = (uint8_t *)xcalloc((size_t)f->n_oct, 1);
  f->node_of = (nptr *)xcalloc((size_t)n, sizeof(nptr));
  f->list0 = (int64_t *)xcalloc((size_t)n + 1, 8);
  /* pass 1: octet keys from the (even, even, even) nodes, in traversal order
   * (which is (z,y,x) order, so already sorted) */
  int64_t no = 0;
  for (pqptr pq = g->pquad; pq; pq = pq->next) {
    long z = pq->z;
    for (cqptr cq0 = pq->loc; cq0 < pq->loc + pq->length; cq0++, z++)
      for (cqptr cq = cq0; cq; cq = cq->next) {
        long y = cq->y;
        for (nqptr nq0 = cq->loc; nq0 < cq->loc + cq->length; nq0++, y++)
          for (nqptr nq = nq0; nq; nq = nq->next) {
            if ((nq->x & 1) || (nq->length & 1)) return -1;
            if ((z & 1) || (y & 1)) continue;
            for (long x = nq->x; x < nq->x + nq->length; x += 2) {
              if (no >= f->n_oct) return -1;
              f->key[no++] = PK(x / 2, y / 2, z / 2);
            }
          }
      }
  }
  if (no != f->n_oct) return -1;
  for (int64_t i = 1; i < no; i++)
    if (f->key[i] <= f->key[i - 1]) qsort(f->key, (size_t)no, 8, cmp_u64), i = no;
  /* pass 2: every node and its coordinates in traversal order, then (in
   * parallel) the node map, the links and the list lengths */
  nptr *nodes = (nptr *)xcalloc((size_t)n, sizeof(nptr));
  uint64_t *nxyz = (uint64_t *)xcalloc((size_t)n, 8);
  int64_t k = 0;
  for (pqptr pq = g->pquad; pq; pq = pq->next) {
    long z = pq->z;
    for (cqptr cq0 = pq->loc; cq0 < pq->loc + pq->length; cq0++, z++)
      for (cqptr cq = cq0; cq; cq = cq->next) {
        long y = cq->y;
        for (nqptr nq0 = cq->loc; nq0 < cq->loc + cq->length; nq0++, y++)
          for (nqptr nq = nq0; nq; nq = nq->next) {
            long x = nq->x;
            for (nptr nd = nq->loc; nd < nq->loc + nq->length; nd++, x++) {
              nodes[k] = nd;
              /* bit 63: the x-run goes on past the next node */
              nxyz[k++] = PK(x, y, z) | (nd + 2 < nq->loc + nq->length ? 1ull << 63 : 0ull);
            }
          }
      }
  }
  int64_t *cnt = f->list0 + 1;
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int64_t i = 0; i < n; i++) {
    const uint64_t w = nxyz[i];
    const int64_t x = (int64_t)(w & 0x1FFFFF), y = (int64_t)((w >> 21) & 0x1FFFFF), z = (int64_t)((w >> 42) & 0x1FFFFF);
    const int64_t o = find_key(f->key, no, PK(x / 2, y / 2, z / 2));
    if (o < 0) {
      bad = 1;
      continue;
    }
    const int64_t idx = 8 * o + (z & 1) * 4 + (y & 1) * 2 + (x & 1);
    if (__atomic_exchange_n(&f->node_of[idx], nodes[i], __ATOMIC_RELAXED)) {
      bad = 1;
      continue;
    }
    if (!(x & 1) && !(y & 1) && !(z & 1) && (w >> 63)) f->link[o] = 1;
    if (!lists) continue;
    int64_t c = 0;
    for (partptr p = nodes[i]->ll; p; p = p->ll) c++;
    cnt[idx] = c;
  }
  free(nodes);
  free(nxyz);
  if (bad) return -1;
  if (!lists) return 0;
  for (int64_t i = 0; i < n; i++) f->list0[i + 1] += f->list0[i];
  f->n_listed = f->list0[n];
  f->list = (uint32_t *)xcalloc((size_t)f->n_listed, 4);
  const partptr P0 = global.fst_part;
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < n; i++) {
    int64_t w = f->list0[i];
    for (partptr p = f->node_of[i]->ll; p; p = p->ll) f->list[w++] = (uint32_t)(p - P0);
  }
  return 0;
}

boolean assign_npart(gridls *cur_grid)
{
  pthread_mutex_lock(&g_lock);
  flat_grid f;
  if (flatten_grid(cur_grid, &f, 1)) {
    pthread_mutex_unlock(&g_lock);
    fatal("assign_npart: grid is not made of 2x2x2 node blocks");
  }
  ensure_particles(0);
  float *dens = (float *)xcalloc((size_t)f.n_node, 4);
  if (ahfgpu_deposit_level(ctx_get(), (int64_t)cur_grid->l1dim, f.key, f.link, f.n_oct, f.list0, f.list,
                           cur_grid->masstopartdens, dens)) {
    pthread_mutex_unlock(&g_lock);
    fatal("assign_npart");
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < f.n_node; i++) f.node_of[i]->dens = dens[i];
  cur_grid->size.no_part = (long unsigned)f.n_listed;
  cur_grid->size.no_nodes = (long unsigned)f.n_node;
  free(dens);
  flat_free(&f);
  pthread_mutex_unlock(&g_lock);
  /* src/libamr_serial/density.c:420-437 */
  if (cur_grid->size.no_nodes < AHF_MIN_NNODES) return 0;
  if (simu.np_limit == 0) return 1;
  if (cur_grid->l1dim == 2 * (long unsigned)simu.NGRID_DOM && simu.Nth_dom < 3.) return 1;
  return ((double)cur_grid->size.no_nodes / (double)cur_grid->size.no_part) > 7.5 ? 0 : 1;
}

/* AHF quads of a level held as octets (calloc'd as c_pquad/c_cquad/c_nquad/
 * c_node would, so free_grid() releases them); node_of receives the node of
 * every octet-major index */
static void build_quads(gridls *g, const uint64_t *key, const uint8_t *link, int64_t n_oct, nptr *node_of)
{
  octet_raster_t r;
  if (octet_raster_build(key, link, n_oct, &r)) fatal("out of host memory");
  /* the pquad chain, and for every node plane its cquad slot and its first
   * cquad, nquad and node (in traversal order) */
  typedef struct {
    cqptr slot;
    int32_t z;
    int64_t ic, iq, nd;
  } plane_t;
  int64_t np = 0;
  for (int64_t k = 0; k < r.n_pq; k++) np += r.pq[2 * k + 1];
  plane_t *pl = (plane_t *)xcalloc((size_t)np + 1, sizeof(plane_t));
  pqptr prev_pq = NULL;
  g->no_pquad = 0;
  int64_t ic = 0, iq = 0, nd = 0, j = 0;
  for (int64_t k = 0; k < r.n_pq; k++) {
    pqptr pq = (pqptr)xcalloc(1, sizeof(pquad));
    if (prev_pq) prev_pq->next = pq; else g->pquad = pq;
    prev_pq = pq;
    g->no_pquad++;
    pq->z = r.pq[2 * k];
    pq->length = r.pq[2 * k + 1];
    pq->loc = (cqptr)xcalloc((size_t)pq->length, sizeof(cquad));
    for (int zz = 0; zz < pq->length; zz++, j++) {
      const int z = pq->z + zz;
      pl[j].slot = pq->loc + zz;
      pl[j].z = z;
      pl[j].ic = ic;
      pl[j].iq = iq;
      pl[j].nd = nd;
      while (ic < r.n_cq && r.cq[3 * ic] == z) ic++;
      while (iq < r.n_nq && r.nq[4 * iq] == z) nd += r.nq[4 * iq++ + 3];
    }
  }
  /* the cquads, nquads and nodes of every plane */
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t jp = 0; jp < np; jp++) {
    const int z = pl[jp].z;
    int64_t jc = pl[jp].ic, jq = pl[jp].iq, jn = pl[jp].nd;
    cqptr prev_cq = NULL;
    while (jc < r.n_cq && r.cq[3 * jc] == z) {
      cqptr cq = prev_cq ? (cqptr)xcalloc(1, sizeof(cquad)) : pl[jp].slot;
      if (prev_cq) prev_cq->next = cq;
      prev_cq = cq;
      cq->y = r.cq[3 * jc + 1];
      cq->length = r.cq[3 * jc + 2];
      cq->loc = (nqptr)xcalloc((size_t)cq->length, sizeof(nquad));
      for (int yy = 0; yy < cq->length; yy++) {
        const int y = cq->y + yy;
        nqptr prev_nq = NULL;
        while (jq < r.n_nq && r.nq[4 * jq] == z && r.nq[4 * jq + 1] == y) {
          nqptr nq = prev_nq ? (nqptr)xcalloc(1, sizeof(nquad)) : cq->loc + yy;
          if (prev_nq) prev_nq->next = nq;
          prev_nq = nq;
          nq->x = r.nq[4 * jq + 2];
          nq->length = r.nq[4 * jq + 3];
          nq->loc = (nptr)xcalloc((size_t)nq->length, sizeof(node));
          for (int xx = 0; xx < nq->length; xx++) node_of[r.order[jn++]] = nq->loc + xx;
          jq++;
        }
      }
      jc++;
    }
  }
  free(pl);
  free(g->pquad_array);
  g->pquad_array = (pqptr *)xcalloc((size_t)g->no_pquad, sizeof(pqptr));
  int64_t i = 0;
  for (pqptr pq = g->pquad; pq; pq = pq->next) g->pquad_array[i++] = pq;
  octet_raster_free(&r);
}

/* densities and final particle lists of device level `lev` into grid g */
static void fill_level(gridls *g, int lev, int build)
{
  ahfgpu_ctx *c = ctx_get();
  int64_t L, n_oct, n_listed;
  double crit, m2pd;
  if (ahfgpu_amr_level_sizes(c, lev, &L, &n_oct, &n_listed, &crit, &m2pd)) fatal("level sizes");
  const int64_t nn = 8 * n_oct;
  uint64_t *key = (uint64_t *)xcalloc((size_t)n_oct, 8);
  uint8_t *link = (uint8_t *)xcalloc((size_t)n_oct, 1);
  float *dens = (float *)xcalloc((size_t)nn, 4);
  int64_t *list0 = (int64_t *)xcalloc((size_t)nn + 1, 8);
  uint32_t *list = (uint32_t *)xcalloc((size_t)n_listed + 1, 4);
  if (ahfgpu_amr_level_copy(c, lev, key, link, dens, list0, list)) fatal("level copy");
  nptr *node_of = (nptr *)xcalloc((size_t)nn, sizeof(nptr));
  if (build) {
    build_quads(g, key, link, n_oct, node_of);
  } else {
    flat_grid f;
    if (flatten_grid(g, &f, 0) || f.n_oct != n_oct) fatal("domain grid does not match the device level");
    memcpy(node_of, f.node_of, (size_t)nn * sizeof(nptr));
    flat_free(&f);
  }
  partptr P = global.fst_part;
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < nn; i++) {
    nptr nd = node_of[i];
    nd->dens = dens[i];
    nd->force.tempdens = 0.0f;
    nd->ll = list0[i] < list0[i + 1] ? P + list[list0[i]] : NULL;
    for (int64_t q = list0[i]; q < list0[i + 1]; q++)
      P[list[q]].ll = q + 1 < list0[i + 1] ? P + list[q + 1] : NULL;
  }
  free(key); free(link); free(dens); free(list0); free(list); free(node_of);
}

boolean gen_AMRhierarchy(gridls **grid_list, int *no_grids)
{
  gridls *fin = *grid_list + *no_grids - 1;
  /* src/libamr_serial/generate_grids.c:307-308 */
  if ((!global.fst_cycle && (long)fin->l1dim == global.fin_l1dim) || (long)fin->l1dim == simu.NGRID_MAX) {
    if (global.fst_cycle) {
      global.fin_l1dim = (long)fin->l1dim;
      global.fst_cycle = 0;
    }
    return 0;
  }
  if ((long)fin->l1dim != simu.NGRID_DOM) fatal("gen_AMRhierarchy: expects the domain grid as the finest grid");
  pthread_mutex_lock(&g_lock);
  ensure_particles(1);
  ahfgpu_amr_params p = {simu.NGRID_DOM, simu.NGRID_MAX, simu.Nth_dom, simu.Nth_ref};
  ahfgpu_amr_info info;
  if (ahfgpu_amr_build(ctx_get(), &p, &info)) {
    pthread_mutex_unlock(&g_lock);
    fatal("gen_AMRhierarchy");
  }
  const int dom = *no_grids - 1, nl = info.n_levels;
  /* AHF reallocates one extra gridls whenever it attempts a refinement that
   * is not kept (src/libamr_serial/generate_grids.c:146-192) */
  const int attempted_extra = ((long)(simu.NGRID_DOM << (nl - 1)) != simu.NGRID_MAX) ? 1 : 0;
  gridls *gl = (gridls *)realloc(*grid_list, (size_t)(*no_grids + nl - 1 + attempted_extra) * sizeof(gridls));
  if (!gl) fatal("gen_AMRhierarchy: realloc of the grid list");
  memset(gl + *no_grids, 0, (size_t)(nl - 1 + attempted_extra) * sizeof(gridls));
  *grid_list = gl;
  global.dom_grid = gl + global.domgrid_no;
  fill_level(gl + dom, 0, 0);
  gl[dom].size.no_nodes = (long unsigned)info.n_nodes[0];
  gl[dom].size.no_part = (long unsigned)info.n_part[0];
  for (int k = 1; k < nl + attempted_extra; k++) {
    gridls *coa = gl + dom + k - 1, *f = gl + dom + k;
    f->masstodens = coa->masstodens * AHF_CRITMULTI;
    f->masstopartdens = coa->masstopartdens * AHF_CRITMULTI;
    f->critdens = simu.Nth_ref * f->masstodens;
    f->timecounter = coa->timecounter;
    f->timestep = coa->timestep / 2;
    f->l1dim = coa->l1dim * 2;
    f->spacing = 1.0 / (double)f->l1dim;
    f->spacing2 = f->spacing * f->spacing;
    f->next = 0;
    coa->next = 1;
    if (k < nl) {
      fill_level(f, k, 1);
      f->size.no_nodes = (long unsigned)info.n_nodes[k];
      f->size.no_part = (long unsigned)info.n_part[k];
    }
  }
  *no_grids += nl - 1;
  global.fin_l1dim = (long)gl[*no_grids - 1].l1dim;
  global.fst_cycle = 0;
  pthread_mutex_unlock(&g_lock);
  return 0;
}

/* --------------------------------------------------------- gridinfo ---- */

/* AHF defines these three in the object ahf_gridinfo replaces
 * (src/libahf/ahf_gridinfo.c:43-45); `ahf` is src/common.c:26's */
SRINDEX *spatialRefIndex = NULL;
int *numIsoRef = NULL;
int totnumIsoRef = 0;
__attribute__((weak)) info_ahf ahf;

/* AHF's cosmology (src/libutility/cosmology.c:47-53, 117-175), called as
 * ahf_gridinfo calls it; weak references, so that the library also links
 * without it */
extern double calc_virial(double a) __attribute__((weak));
extern double calc_rho_vir(double a) __attribute__((weak));

void ahf_gridinfo(gridls *grid_list, int curgrid_no)
{
  (void)grid_list;
  if (!calc_virial || !calc_rho_vir) fatal("ahf_gridinfo: AHF's calc_virial and calc_rho_vir are not linked");
  /* src/libahf/ahf_gridinfo.c:126-166 */
  const double a = global.a;
  ahfgpu_gridinfo_params p;
  p.ovlim = calc_virial(a);
  p.rho_vir = (a * a * a) * calc_rho_vir(a);
  p.refine_mass = simu.pmass * simu.med_weight;
  p.boxsize = simu.boxsize;
  pthread_mutex_lock(&g_lock);
  ahfgpu_ctx *c = ctx_get();
  int64_t L, n_oct, n_listed;
  double crit, m2pd;
  const int nref = curgrid_no - global.domgrid_no;
  /* the grids must be the library's last hierarchy: as many refinements,
   * and the finest one of the expected size */
  if (nref < 0 || ahfgpu_amr_level_sizes(c, nref, &L, &n_oct, &n_listed, &crit, &m2pd) ||
      (uint64_t)L != (uint64_t)global.dom_grid[nref].l1dim ||
      ahfgpu_amr_level_sizes(c, nref + 1, &L, &n_oct, &n_listed, &crit, &m2pd) == 0) {
    pthread_mutex_unlock(&g_lock);
    fatal("ahf_gridinfo: the grids are not the hierarchy gen_AMRhierarchy built last");
  }
  ahfgpu_gridinfo_info info;
  if (ahfgpu_gridinfo(c, &p, &info)) {
    pthread_mutex_unlock(&g_lock);
    fatal("ahf_gridinfo");
  }
  ahf.min_ref = info.min_ref;
  ahf.no_grids = info.no_grids;
  global.max_ovdens = info.max_ovdens;
  /* src/libahf/ahf_gridinfo.c:720-763: one entry per colour started (names
   * 1..n_colours; entry 0 and the merged colours keep -1, -1 and no
   * periodicity; the last entry stays zero, as the reference leaves it) */
  const int ncol = info.n_colours;
  spatialRefIndex = (SRINDEX *)xcalloc((size_t)ncol + 1, sizeof(SRINDEX));
  for (int i = 0; i < ncol; i++) spatialRefIndex[i].refLevel = spatialRefIndex[i].isoRefIndex = -1;
  numIsoRef = (int *)xcalloc(info.no_grids > 0 ? (size_t)info.no_grids : 1, sizeof(int));
  totnumIsoRef = info.tot_iso;
  for (int g = 0; g < info.no_grids; g++) {
    gridls *cg = global.dom_grid + info.min_ref + g;
    flat_grid f;
    if (flatten_grid(cg, &f, 0)) {
      pthread_mutex_unlock(&g_lock);
      fatal("ahf_gridinfo: a grid is not made of octets");
    }
    numIsoRef[g] = info.num_iso[g];
    int32_t *col = (int32_t *)xcalloc((size_t)f.n_node, 4);
    int32_t *refs = (int32_t *)xcalloc(6 * (size_t)(info.num_iso[g] > 0 ? info.num_iso[g] : 1), 4);
    if (ahfgpu_amr_level_sizes(c, info.min_ref + g, &L, &n_oct, &n_listed, &crit, &m2pd) || n_oct != f.n_oct ||
        ahfgpu_gridinfo_results(c, g, col, refs)) {
      pthread_mutex_unlock(&g_lock);
      fatal("ahf_gridinfo: grid results");
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < f.n_node; i++) f.node_of[i]->force.colour = col[i];
    for (int k = 0; k < info.num_iso[g]; k++) {
      const int32_t *e = refs + 6 * k;
      if (e[0] < 1 || e[0] > ncol) {
        pthread_mutex_unlock(&g_lock);
        fatal("ahf_gridinfo: a refinement's colour is out of range");
      }
      SRINDEX *r = spatialRefIndex + e[0];
      r->refLevel = e[1];
      r->isoRefIndex = e[2];
      r->periodic.x = e[3];
      r->periodic.y = e[4];
      r->periodic.z = e[5];
    }
    free(col);
    free(refs);
    flat_free(&f);
  }
  pthread_mutex_unlock(&g_lock);
}

/* ----------------------------------------------------------- haloes ---- */

static int run_halos(HALO **h, long unsigned nh, int mode, int clear_phi0)
{
  ahfgpu_ctx *c = ctx_get();
  ensure_particles(0);
  uint64_t tot = 0;
  for (long unsigned i = 0; i < nh; i++) tot += h[i]->npart;
  double *cen = (double *)xcalloc(3 * nh, 8);
  uint64_t *np = (uint64_t *)xcalloc(nh, 8), *ip = (uint64_t *)xcalloc(tot, 8);
  uint64_t w = 0;
  for (long unsigned i = 0; i < nh; i++) {
    cen[3 * i] = h[i]->pos.x;
    cen[3 * i + 1] = h[i]->pos.y;
    cen[3 * i + 2] = h[i]->pos.z;
    np[i] = h[i]->npart;
    for (uint64_t j = 0; j < h[i]->npart; j++) ip[w++] = h[i]->ipart[j];
  }
  ahfgpu_unbind_params up = {simu.AHF_MINPART, 0, simu.AHF_VTUNE, r_fac, v_fac, phi_fac, Hubble,
                             global.ovlim, global.rho_vir, rho_fac};
  int rc = ahfgpu_set_halos(c, cen, np, ip, nh);
  if (!rc) rc = ahfgpu_unbind(c, &up, mode, 0.0, NULL);
  double *props = (double *)xcalloc(AHFGPU_PROPS * nh, 8);
  if (!rc) rc = ahfgpu_unbind_results(c, np, props, ip);
  if (!rc) {
    w = 0;
    for (long unsigned i = 0; i < nh; i++) {
      const uint64_t n_in = h[i]->npart;
      const double *q = props + AHFGPU_PROPS * i;
      int set = (int)q[5];
      if (!clear_phi0) set &= ~AHFGPU_SET_PHI0;
      if (mode & AHFGPU_UNBIND_SORT_ONLY) {
        /* sort_halo_particles reorders the array in place (below NminPerHalo
         * the order is kept, so this copy changes nothing) */
        memcpy(h[i]->ipart, ip + w, n_in * 8);
      } else if (set) {
        /* rem_unbound installs a new array, rem_outsideRvir realloc()s the
         * old one: either way the caller owns a malloc'd array of npart */
        long unsigned *lst = (long unsigned *)realloc(h[i]->ipart, (np[i] ? np[i] : 1) * 8);
        if (!lst) fatal("out of host memory");
        memcpy(lst, ip + w, np[i] * 8);
        h[i]->ipart = lst;
        h[i]->npart = np[i];
        if (set & AHFGPU_SET_MVIR) h[i]->M_vir = q[0];
        if (set & AHFGPU_SET_RVIR) h[i]->R_vir = q[1];
        if (set & AHFGPU_SET_PHI0) h[i]->Phi0 = q[2];
        if (set & AHFGPU_SET_OVDENS) h[i]->ovdens = q[3];
      }
      w += n_in;
    }
  }
  ahfgpu_release_halos(c);
  free(cen); free(np); free(ip); free(props);
  return rc;
}

static int run_profiles(HALO **h, long unsigned nh);
static int gather_halos(HALO **h, long unsigned nh, int construct);

/* the batched functions take the haloes as an array of pointers */
static HALO **halo_ptrs(HALO *h, long unsigned nh)
{
  HALO **hp = (HALO **)xcalloc(nh, sizeof(HALO *));
  for (long unsigned i = 0; i < nh; i++) hp[i] = h + i;
  return hp;
}

/*
 * The per-halo functions are called from AHF's OpenMP halo loop
 * (src/libahf/ahf_halos.c:504-510 around ahf_halos_sfc_constructHalo), one
 * halo per call.  Calls made at the same time are combined: each call queues
 * its request and waits; whichever caller finds no batch running takes
 * every queued request and runs it as one batched pass per kind of request
 * (all sort_halo_particles together, all rem_unbound together, ...), then
 * wakes the callers whose requests it ran.  A thread's requests for one halo
 * follow each other (each call returns before the next is made), so a batch
 * holds each halo at most once and the haloes in it are independent.
 */
enum { REQ_SORT, REQ_UNBIND, REQ_RVIR0, REQ_RVIR1, REQ_PROFILES, REQ_GATHER, REQ_KINDS };

typedef struct halo_req {
  HALO *h;
  int kind, rc, done;
  struct halo_req *next;
} halo_req;

static pthread_mutex_t q_lock = PTHREAD_MUTEX_INITIALIZER;
static pthread_cond_t q_cv = PTHREAD_COND_INITIALIZER;
static halo_req *q_head = NULL, **q_tail = &q_head;
/* q_running and q_len change under q_lock and are also read without it
 * (atomically) by threads that spin */
static int q_running = 0, q_len = 0;
static uint64_t q_batches = 0, q_requests = 0;
/* how long a new batch waits for the rest of the caller's OpenMP team to
 * queue its next request (each thread has at most one request queued, so
 * the team size bounds a batch): until no request has arrived for
 * GATHER_IDLE_US, at most GATHER_MAX_US.  A thread whose request is in a
 * running batch spins on its own done flag for up to WAIT_SPIN_US (a pass
 * takes a few hundred microseconds; a woken sleeper takes tens to run
 * again), then sleeps until the batch ends. */
#define GATHER_IDLE_US 40
#define GATHER_MAX_US 200
#define WAIT_SPIN_US 4000

static int run_kind(int kind, HALO **hp, long unsigned n)
{
  switch (kind) {
  case REQ_SORT: return run_halos(hp, n, AHFGPU_UNBIND_BATCHED | AHFGPU_UNBIND_SORT_ONLY, 0);
  case REQ_UNBIND: return run_halos(hp, n, AHFGPU_UNBIND_BATCHED | AHFGPU_UNBIND_NO_SORT, 1);
  case REQ_RVIR0: return run_halos(hp, n, AHFGPU_UNBIND_BATCHED | AHFGPU_UNBIND_RVIR_ONLY, 1);
  case REQ_RVIR1: return run_halos(hp, n, AHFGPU_UNBIND_BATCHED | AHFGPU_UNBIND_RVIR_ONLY, 0);
  case REQ_PROFILES: return run_profiles(hp, n);
  default: return gather_halos(hp, n, 0);
  }
}

/* every request of the batch, one batched pass per kind */
static void run_batch(halo_req *batch)
{
  long unsigned n[REQ_KINDS] = {0}, tot = 0;
  for (halo_req *r = batch; r; r = r->next) n[r->kind]++, tot++;
  HALO **hp = (HALO **)xcalloc(tot, sizeof(HALO *));
  pthread_mutex_lock(&g_lock);
  for (int k = 0; k < REQ_KINDS; k++) {
    if (!n[k]) continue;
    long unsigned m = 0;
    for (halo_req *r = batch; r; r = r->next)
      if (r->kind == k) hp[m++] = r->h;
    const int rc = run_kind(k, hp, m);
    for (halo_req *r = batch; r; r = r->next)
      if (r->kind == k) r->rc = rc;
  }
  q_batches++;
  q_requests += tot;
  pthread_mutex_unlock(&g_lock);
  free(hp);
}

static inline void cpu_relax(void)
{
#if defined(__x86_64__) || defined(__i386__)
  __builtin_ia32_pause();
#endif
}

static int64_t mono_us(void)
{
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (int64_t)t.tv_sec * 1000000 + t.tv_nsec / 1000;
}

static int submit(HALO *h, int kind)
{
  halo_req req = {h, kind, 0, 0, NULL};
#ifdef _OPENMP
  const int team = omp_in_parallel() ? omp_get_num_threads() : 1;
#else
  const int team = 1;
#endif
  pthread_mutex_lock(&q_lock);
  *q_tail = &req;
  q_tail = &req.next;
  __atomic_store_n(&q_len, q_len + 1, __ATOMIC_RELEASE);
  while (!__atomic_load_n(&req.done, __ATOMIC_ACQUIRE)) {
    if (q_running) {
      /* a batch runs: spin on this request without the lock until it is
       * done or the batch ends, then sleep if the spin ran out */
      pthread_mutex_unlock(&q_lock);
      const int64_t t0 = mono_us();
      int spun_out = 0;
      for (unsigned k = 1; !__atomic_load_n(&req.done, __ATOMIC_ACQUIRE) &&
                           __atomic_load_n(&q_running, __ATOMIC_ACQUIRE);
           k++) {
        cpu_relax();
        if ((k & 255) == 0 && mono_us() - t0 > WAIT_SPIN_US) {
          spun_out = 1;
          break;
        }
      }
      pthread_mutex_lock(&q_lock);
      while (spun_out && q_running && !__atomic_load_n(&req.done, __ATOMIC_ACQUIRE))
        pthread_cond_wait(&q_cv, &q_lock);
      continue;
    }
    /* no batch running: this thread runs the next one.  It first lets the
     * rest of its team queue their requests, then takes every queued
     * request, this one included */
    __atomic_store_n(&q_running, 1, __ATOMIC_RELEASE);
    if (team > 1 && q_len < team) {
      pthread_mutex_unlock(&q_lock);
      const int64_t t0 = mono_us();
      int64_t last = t0;
      int seen = __atomic_load_n(&q_len, __ATOMIC_ACQUIRE);
      for (;;) {
        const int len = __atomic_load_n(&q_len, __ATOMIC_ACQUIRE);
        if (len >= team) break;
        const int64_t now = mono_us();
        if (len != seen) {
          seen = len;
          last = now;
        } else if (now - last > GATHER_IDLE_US || now - t0 > GATHER_MAX_US) {
          break;
        }
        cpu_relax();
      }
      pthread_mutex_lock(&q_lock);
    }
    halo_req *batch = q_head;
    q_head = NULL;
    q_tail = &q_head;
    __atomic_store_n(&q_len, 0, __ATOMIC_RELEASE);
    pthread_mutex_unlock(&q_lock);
    run_batch(batch);
    pthread_mutex_lock(&q_lock);
    for (halo_req *r = batch, *nx; r; r = nx) {
      nx = r->next; /* r lives on its caller's stack until done is seen */
      __atomic_store_n(&r->done, 1, __ATOMIC_RELEASE);
    }
    __atomic_store_n(&q_running, 0, __ATOMIC_RELEASE);
    pthread_cond_broadcast(&q_cv);
  }
  pthread_mutex_unlock(&q_lock);
  return req.rc;
}

void ahf_gpu_halo_call_stats(uint64_t *batches, uint64_t *requests)
{
  pthread_mutex_lock(&g_lock);
  if (batches) *batches = q_batches;
  if (requests) *requests = q_requests;
  pthread_mutex_unlock(&g_lock);
}

void sort_halo_particles(HALO *halo)
{
  if (halo->npart < (long unsigned)simu.AHF_MINPART) return;
  if (submit(halo, REQ_SORT)) fatal("sort_halo_particles");
}

void rem_unbound(HALO *halo)
{
  if (halo->npart < (long unsigned)simu.AHF_MINPART) return;
  if (submit(halo, REQ_UNBIND)) fatal("rem_unbound");
}

int ahf_unbind_halos(HALO *halos, long unsigned nhalos)
{
  if (nhalos == 0) return 0;
  HALO **hp = halo_ptrs(halos, nhalos);
  pthread_mutex_lock(&g_lock);
  int rc = run_halos(hp, nhalos, AHFGPU_UNBIND_BATCHED, 1);
  pthread_mutex_unlock(&g_lock);
  free(hp);
  return rc;
}

void rem_outsideRvir(HALO *halo, int icall)
{
  if (halo->npart < (long unsigned)simu.AHF_MINPART) return;
  if (submit(halo, icall == 0 ? REQ_RVIR0 : REQ_RVIR1)) fatal("rem_outsideRvir");
}

int ahf_construct_halos(HALO *halos, long unsigned nhalos)
{
  if (nhalos == 0) return 0;
  HALO **hp = halo_ptrs(halos, nhalos);
  pthread_mutex_lock(&g_lock);
  int rc = run_halos(hp, nhalos, AHFGPU_UNBIND_BATCHED | AHFGPU_UNBIND_WITH_RVIR, 1);
  pthread_mutex_unlock(&g_lock);
  free(hp);
  return rc;
}

/* --------------------------------------------------------- profiles ---- */

/* c_profile (src/libutility/alloc_struct.c:614-640): the profile arrays of
 * nbins entries, each calloc()ed on its own as AHF frees them */
static void alloc_profile(HALO *h, int nbins)
{
  HALOPROFILE *q = &h->prof;
  q->nbins = nbins;
  double **arr[] = {&q->r, &q->nvpart, &q->ovdens, &q->dens, &q->v2_circ, &q->v_esc2, &q->sig_v, &q->Ekin,
                    &q->Epot, &q->Lx, &q->Ly, &q->Lz, &q->axis1, &q->E1x, &q->E1y, &q->E1z, &q->axis2,
                    &q->E2x, &q->E2y, &q->E2z, &q->axis3, &q->E3x, &q->E3y, &q->E3z};
  for (size_t k = 0; k < sizeof(arr) / sizeof(arr[0]); k++) *arr[k] = (double *)xcalloc((size_t)nbins, 8);
  q->npart = (long unsigned *)xcalloc((size_t)nbins, sizeof(long unsigned));
}

/* the fields HaloProfiles sets (src/libahf/ahf_halos.c:4069, 4674-4710,
 * 4922-5018) from one halo's AHFGPU_PROF_NP properties and nbins x
 * AHFGPU_PROF_NQ profile values */
static void store_profile(HALO *h, int nbins, const double *P, const double *Q)
{
  alloc_profile(h, nbins);
  HALOPROFILE *q = &h->prof;
  for (int b = 0; b < nbins; b++) {
    const double *v = Q + (size_t)b * AHFGPU_PROF_NQ;
    q->r[b] = v[0];
    q->npart[b] = (long unsigned)v[1];
    q->nvpart[b] = v[2];
    q->ovdens[b] = v[3];
    q->dens[b] = v[4];
    q->v2_circ[b] = v[5];
    q->v_esc2[b] = v[6];
    q->sig_v[b] = v[7];
    q->Ekin[b] = v[8];
    q->Epot[b] = v[9];
    q->Lx[b] = v[10];
    q->Ly[b] = v[11];
    q->Lz[b] = v[12];
    q->E1x[b] = v[13];
    q->E1y[b] = v[14];
    q->E1z[b] = v[15];
    q->E2x[b] = v[16];
    q->E2y[b] = v[17];
    q->E2z[b] = v[18];
    q->E3x[b] = v[19];
    q->E3y[b] = v[20];
    q->E3z[b] = v[21];
    q->axis1[b] = v[22];
    q->axis2[b] = v[23];
    q->axis3[b] = v[24];
  }
  h->M_vir = P[1];
  h->vel.x = P[2];
  h->vel.y = P[3];
  h->vel.z = P[4];
  h->R_max = P[5];
  h->V2_max = P[6];
  h->r2 = P[7];
  h->sigV = P[8];
  h->v_esc2 = P[9];
  h->Ekin = P[10];
  h->Epot = P[11];
  h->axis.x = P[12];
  h->axis.y = P[13];
  h->axis.z = P[14];
  h->E1.x = P[15];
  h->E1.y = P[16];
  h->E1.z = P[17];
  h->E2.x = P[18];
  h->E2.y = P[19];
  h->E2.z = P[20];
  h->E3.x = P[21];
  h->E3.y = P[22];
  h->E3.z = P[23];
  h->AngMom.x = P[24];
  h->AngMom.y = P[25];
  h->AngMom.z = P[26];
  h->lambda = P[27];
  h->lambdaE = P[28];
  h->cNFW = P[29];
  h->R1 = P[30];
  h->cR1 = P[31];
  h->SurfP = P[32];
  h->fMhires = P[33];
  h->mbp_offset = P[34];
  h->pos_mbp.x = P[35];
  h->pos_mbp.y = P[36];
  h->pos_mbp.z = P[37];
  h->vel_mbp.x = P[38];
  h->vel_mbp.y = P[39];
  h->vel_mbp.z = P[40];
  h->com_offset = P[41];
  h->pos_com.x = P[42];
  h->pos_com.y = P[43];
  h->pos_com.z = P[44];
}

static ahfgpu_profile_params profile_params(void)
{
  ahfgpu_profile_params pp = {simu.AHF_MINPART, 0, r_fac, v_fac, m_fac, phi_fac, Hubble};
  return pp;
}

/* HaloProfiles of the resident haloes (last unbinding result, or the lists
 * as given with phi_rvir) into h[0..nh) */
static int profiles_into(ahfgpu_ctx *c, HALO **h, long unsigned nh, const double *phi_rvir)
{
  const ahfgpu_profile_params pp = profile_params();
  int32_t *nb = (int32_t *)xcalloc(nh, 4);
  int rc = ahfgpu_halo_profiles(c, &pp, phi_rvir, nb);
  uint64_t tot = 0;
  for (long unsigned i = 0; i < nh; i++) tot += (uint64_t)nb[i];
  double *P = (double *)xcalloc(nh * AHFGPU_PROF_NP, 8), *Q = (double *)xcalloc(tot * AHFGPU_PROF_NQ + 1, 8);
  if (!rc) rc = ahfgpu_profile_results(c, P, Q);
  if (!rc) {
    uint64_t w = 0;
    for (long unsigned i = 0; i < nh; i++) {
      if (nb[i] <= 0) continue;
      store_profile(h[i], nb[i], P + (size_t)i * AHFGPU_PROF_NP, Q + w * AHFGPU_PROF_NQ);
      w += (uint64_t)nb[i];
    }
  }
  free(nb); free(P); free(Q);
  return rc;
}

static int run_profiles(HALO **h, long unsigned nh)
{
  ahfgpu_ctx *c = ctx_get();
  ensure_particles(0);
  uint64_t tot = 0;
  for (long unsigned i = 0; i < nh; i++) tot += h[i]->npart;
  double *cen = (double *)xcalloc(5 * nh, 8), *pr = cen + 3 * nh;
  uint64_t *np = (uint64_t *)xcalloc(nh, 8), *ip = (uint64_t *)xcalloc(tot + 1, 8);
  uint64_t w = 0;
  for (long unsigned i = 0; i < nh; i++) {
    cen[3 * i] = h[i]->pos.x;
    cen[3 * i + 1] = h[i]->pos.y;
    cen[3 * i + 2] = h[i]->pos.z;
    pr[2 * i] = h[i]->Phi0;
    pr[2 * i + 1] = h[i]->R_vir;
    np[i] = h[i]->npart;
    for (uint64_t j = 0; j < h[i]->npart; j++) ip[w++] = h[i]->ipart[j];
  }
  int rc = ahfgpu_set_halos(c, cen, np, ip, nh);
  if (!rc) rc = profiles_into(c, h, nh, pr);
  ahfgpu_release_halos(c);
  free(cen); free(np); free(ip);
  return rc;
}

int HaloProfiles(HALO *halo)
{
  if (halo->npart < (long unsigned)simu.AHF_MINPART) return 1;
  if (submit(halo, REQ_PROFILES)) fatal("HaloProfiles");
  return 1;
}

int ahf_halo_profiles(HALO *halos, long unsigned nhalos)
{
  if (nhalos == 0) return 0;
  HALO **hp = halo_ptrs(halos, nhalos);
  pthread_mutex_lock(&g_lock);
  int rc = run_profiles(hp, nhalos);
  pthread_mutex_unlock(&g_lock);
  free(hp);
  return rc;
}

/* the halo gather for nh haloes; with construct, the steps after it too
 * (construct == 2: and HaloProfiles, on the lists the device holds).
 * Haloes whose npart is 0 are skipped (src/libahf/ahf_halos_sfc.c:122-123
 * for the construction; the gather alone appends nothing for them) */
static int gather_halos(HALO **h, long unsigned nh, int construct)
{
  ahfgpu_ctx *c = ctx_get();
  ensure_keys();
  double *cen = (double *)xcalloc(4 * nh, 8), *rad = cen + 3 * nh;
  uint8_t *skip = (uint8_t *)xcalloc(nh, 1);
  for (long unsigned i = 0; i < nh; i++) {
    cen[3 * i] = h[i]->pos.x;
    cen[3 * i + 1] = h[i]->pos.y;
    cen[3 * i + 2] = h[i]->pos.z;
    rad[i] = h[i]->gatherRad;
    skip[i] = construct && h[i]->npart == 0;
  }
  uint64_t *np = (uint64_t *)xcalloc(nh, 8);
  int rc = ahfgpu_gather_halos(c, cen, rad, skip, nh);
  if (!rc) rc = ahfgpu_halo_members(c, np, NULL);
  uint64_t tot = 0;
  for (long unsigned i = 0; i < nh; i++) tot += np[i];
  uint64_t *ip = (uint64_t *)xcalloc(tot, 8);
  double *props = (double *)xcalloc(AHFGPU_PROPS * nh, 8);
  if (!rc && construct) {
    ahfgpu_unbind_params up = {simu.AHF_MINPART, 0, simu.AHF_VTUNE, r_fac, v_fac, phi_fac, Hubble,
                               global.ovlim, global.rho_vir, rho_fac};
    rc = ahfgpu_unbind(c, &up, AHFGPU_UNBIND_BATCHED | AHFGPU_UNBIND_WITH_RVIR, 0.0, NULL);
    uint64_t *nin = (uint64_t *)xcalloc(nh, 8);
    memcpy(nin, np, nh * 8);
    if (!rc) rc = ahfgpu_unbind_results(c, np, props, ip);
    /* results sit at the offsets of the gathered lists */
    uint64_t w = 0, r = 0;
    for (long unsigned i = 0; !rc && i < nh; i++) {
      memmove(ip + w, ip + r, np[i] * 8);
      w += np[i];
      r += nin[i];
    }
    free(nin);
  } else if (!rc) {
    rc = ahfgpu_halo_members(c, NULL, ip);
  }
  if (!rc) {
    uint64_t w = 0;
    for (long unsigned i = 0; i < nh; i++) {
      if (skip[i]) continue;
      if (construct) {
        /* ahf_halos_sfc_constructHalo starts from an empty list */
        h[i]->nll = 0;
        h[i]->ll = NULL;
        h[i]->npart = 0;
        h[i]->ipart = NULL;
      }
      if (np[i]) {
        long unsigned *lst = (long unsigned *)realloc(h[i]->ipart, (h[i]->npart + np[i]) * 8);
        if (!lst) fatal("out of host memory");
        memcpy(lst + h[i]->npart, ip + w, np[i] * 8);
        h[i]->ipart = lst;
        h[i]->npart += np[i];
      }
      w += np[i];
      if (construct) {
        const double *q = props + AHFGPU_PROPS * i;
        const int set = (int)q[5];
        if (set & AHFGPU_SET_MVIR) h[i]->M_vir = q[0];
        if (set & AHFGPU_SET_RVIR) h[i]->R_vir = q[1];
        if (set & AHFGPU_SET_PHI0) h[i]->Phi0 = q[2];
        if (set & AHFGPU_SET_OVDENS) h[i]->ovdens = q[3];
      }
    }
  }
  if (!rc && construct == 2) {
    /* skipped haloes have no result: their npart of 0 keeps them below
     * minpart, as AHF never profiles them either */
    rc = profiles_into(c, h, nh, NULL);
  }
  ahfgpu_release_halos(c);
  free(cen); free(skip); free(np); free(ip); free(props);
  return rc;
}

void ahf_halos_sfc_gatherParts(HALO *halo)
{
  if (submit(halo, REQ_GATHER)) fatal("ahf_halos_sfc_gatherParts");
}

int ahf_gather_construct_halos(HALO *halos, long unsigned nhalos)
{
  if (nhalos == 0) return 0;
  HALO **hp = halo_ptrs(halos, nhalos);
  pthread_mutex_lock(&g_lock);
  int rc = gather_halos(hp, nhalos, 1);
  pthread_mutex_unlock(&g_lock);
  free(hp);
  return rc;
}

int ahf_construct_profile_halos(HALO *halos, long unsigned nhalos)
{
  if (nhalos == 0) return 0;
  HALO **hp = halo_ptrs(halos, nhalos);
  pthread_mutex_lock(&g_lock);
  int rc = gather_halos(hp, nhalos, 2);
  pthread_mutex_unlock(&g_lock);
  free(hp);
  return rc;
}
