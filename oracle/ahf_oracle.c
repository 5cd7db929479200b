/*
 * ahf_oracle.c -- CPU restatement of AHF's particle->grid deposit,
 * refinement and unbinding path.  TEST INFRASTRUCTURE ONLY (the checker for
 * the HIP path); never linked into the product.
 *
 * Grids are held as explicit block lists with the same meaning as AHF's
 * quads (src/tdef.h:144-183): a grid is a list of pquads (runs of planes),
 * every plane a list of cquads (runs of rows), every row a list of nquads
 * (runs of nodes).  Nodes are numbered in AHF's traversal order (raster
 * z, y, x), particle lists are singly linked through next[] with AHF's head
 * insertion, and densities are FP32 accumulated as float += double.  All
 * neighbour lookups follow the block rules of src/libamr_serial/get_nnodes.c
 * (a neighbour is found only inside the same block or across the periodic
 * wrap), which is what makes the result bit-identical to the reference.
 */
#define _POSIX_C_SOURCE 200809L
#include "ahf_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../ahf_amd/csrc/snapfile.h"
#include "oracle_profiles.h"

_Static_assert(ORACLE_PROF_NQ == AHF_PROF_NQ && ORACLE_PROF_NP == AHF_PROF_NP, "profile value counts");

#define MIN_NNODES 125     /* src/param.h:54 */
#define CRITMULTI 8.0      /* src/param.h:118 */
#define MACHINE_ZERO 5e-16 /* src/param.h:125 */
#define PI_AHF 3.14159265358979323846264338 /* src/param.h:143 */

/* ------------------------------------------------------ small vectors ---- */

typedef struct { int64_t *v; int64_t n, cap; } vec64;

static void v_push(vec64 *a, int64_t x)
{
  if (a->n == a->cap) {
    a->cap = a->cap ? 2 * a->cap : 64;
    a->v = (int64_t *)realloc(a->v, (size_t)a->cap * sizeof(int64_t));
    if (!a->v) { fprintf(stderr, "oracle: out of memory\n"); exit(3); }
  }
  a->v[a->n++] = x;
}
static void v_free(vec64 *a) { free(a->v); memset(a, 0, sizeof(*a)); }

/* --------------------------------------------------------------- grid ---- */

typedef struct {
  int64_t L;
  double critdens, masstodens, masstopartdens;
  uint64_t size_no_part, size_no_nodes;
  /* pquads: z0, len, first plane */
  vec64 pq_z0, pq_len, pq_pl0;
  /* planes: z, pquad, first cquad */
  vec64 pl_z, pl_pq, pl_cq0;
  /* cquads: y0, len, first row, plane */
  vec64 cq_y0, cq_len, cq_rw0, cq_pl;
  /* rows: y, cquad, plane, first nquad */
  vec64 rw_y, rw_cq, rw_pl, rw_nq0;
  /* nquads: x0, len, row, first node */
  vec64 nq_x0, nq_len, nq_rw, nq_nd0;
  /* nodes */
  int64_t nnode;
  int64_t *nd_nq;   /* nquad of node */
  float *dens;
  int64_t *head;    /* particle list head, -1 = empty */
  int64_t anomalies;
} grid_t;

static void grid_free(grid_t *g)
{
  vec64 *vs[] = {&g->pq_z0, &g->pq_len, &g->pq_pl0, &g->pl_z, &g->pl_pq, &g->pl_cq0,
                 &g->cq_y0, &g->cq_len, &g->cq_rw0, &g->cq_pl, &g->rw_y, &g->rw_cq,
                 &g->rw_pl, &g->rw_nq0, &g->nq_x0, &g->nq_len, &g->nq_rw, &g->nq_nd0};
  for (size_t i = 0; i < sizeof(vs) / sizeof(vs[0]); i++) v_free(vs[i]);
  free(g->nd_nq); free(g->dens); free(g->head);
  memset(g, 0, sizeof(*g));
}

/* builder: blocks are appended in traversal order */
static void b_pquad(grid_t *g, int64_t z0)
{
  v_push(&g->pq_z0, z0); v_push(&g->pq_len, 0); v_push(&g->pq_pl0, g->pl_z.n);
}
static void b_plane(grid_t *g, int64_t z)
{
  v_push(&g->pl_z, z); v_push(&g->pl_pq, g->pq_z0.n - 1); v_push(&g->pl_cq0, g->cq_y0.n);
  g->pq_len.v[g->pq_len.n - 1]++;
}
static void b_cquad(grid_t *g, int64_t y0)
{
  v_push(&g->cq_y0, y0); v_push(&g->cq_len, 0); v_push(&g->cq_rw0, g->rw_y.n);
  v_push(&g->cq_pl, g->pl_z.n - 1);
}
static void b_row(grid_t *g, int64_t y)
{
  v_push(&g->rw_y, y); v_push(&g->rw_cq, g->cq_y0.n - 1); v_push(&g->rw_pl, g->pl_z.n - 1);
  v_push(&g->rw_nq0, g->nq_x0.n);
  g->cq_len.v[g->cq_len.n - 1]++;
}
static void b_nquad(grid_t *g, int64_t x0, int64_t len)
{
  v_push(&g->nq_x0, x0); v_push(&g->nq_len, len); v_push(&g->nq_rw, g->rw_y.n - 1);
  v_push(&g->nq_nd0, g->nnode);
  g->nnode += len;
}
/* close the structure: sentinels and node arrays */
static void b_finish(grid_t *g)
{
  v_push(&g->pq_pl0, g->pl_z.n);
  v_push(&g->pl_cq0, g->cq_y0.n);
  v_push(&g->cq_rw0, g->rw_y.n);
  v_push(&g->rw_nq0, g->nq_x0.n);
  v_push(&g->nq_nd0, g->nnode);
  int64_t n = g->nnode ? g->nnode : 1;
  g->nd_nq = (int64_t *)malloc((size_t)n * sizeof(int64_t));
  g->dens = (float *)calloc((size_t)n, sizeof(float));
  g->head = (int64_t *)malloc((size_t)n * sizeof(int64_t));
  for (int64_t q = 0; q < g->nq_x0.n; q++)
    for (int64_t i = 0; i < g->nq_len.v[q]; i++) g->nd_nq[g->nq_nd0.v[q] + i] = q;
  for (int64_t i = 0; i < g->nnode; i++) g->head[i] = -1;
}

#define NPQ(g) ((g)->pq_z0.n)
#define NPL(g) ((g)->pl_z.n)
#define NCQ(g) ((g)->cq_y0.n)
#define NRW(g) ((g)->rw_y.n)
#define NNQ(g) ((g)->nq_x0.n)

/* coordinates of a node */
static void node_xyz(const grid_t *g, int64_t n, int64_t *x, int64_t *y, int64_t *z)
{
  int64_t q = g->nd_nq[n], r = g->nq_rw.v[q];
  *x = g->nq_x0.v[q] + (n - g->nq_nd0.v[q]);
  *y = g->rw_y.v[r];
  *z = g->pl_z.v[g->rw_pl.v[r]];
}

/* plane index holding z, or -1 (searches every pquad, as the z-wrap does) */
static int64_t find_plane(const grid_t *g, int64_t z)
{
  int64_t lo = 0, hi = NPL(g) - 1;
  while (lo <= hi) {
    int64_t m = (lo + hi) / 2;
    if (g->pl_z.v[m] == z) return m;
    if (g->pl_z.v[m] < z) lo = m + 1; else hi = m - 1;
  }
  return -1;
}

/* row index of row y on plane p (searching the plane's cquad chain), or -1 */
static int64_t find_row(const grid_t *g, int64_t p, int64_t y)
{
  int64_t lo = g->cq_rw0.v[g->pl_cq0.v[p]], hi = g->cq_rw0.v[g->pl_cq0.v[p + 1]] - 1;
  while (lo <= hi) {
    int64_t m = (lo + hi) / 2;
    if (g->rw_y.v[m] == y) return m;
    if (g->rw_y.v[m] < y) lo = m + 1; else hi = m - 1;
  }
  return -1;
}

/* node at x on row r (searching the row's nquad chain), or -1 */
static int64_t find_node(const grid_t *g, int64_t r, int64_t x)
{
  int64_t lo = g->rw_nq0.v[r], hi = g->rw_nq0.v[r + 1] - 1, best = -1;
  while (lo <= hi) {
    int64_t m = (lo + hi) / 2;
    if (g->nq_x0.v[m] <= x) { best = m; lo = m + 1; } else hi = m - 1;
  }
  if (best < 0) return -1;
  if (x < g->nq_x0.v[best] + g->nq_len.v[best]) return g->nq_nd0.v[best] + (x - g->nq_x0.v[best]);
  return -1;
}

/*
 * tsc: the 27 TSC neighbours of node s, tsc[(dz+1)*9+(dy+1)*3+(dx+1)], -1 if
 * AHF would not find them: get_TSCnodes/search_TSCplane/search_TSCcolumn,
 * src/libamr_serial/get_nnodes.c:459-853.
 *  - plane z+-1: same pquad, or across the periodic wrap (:700-780, :795-850);
 *    if the centre node (x,y,z+-1) is absent the whole plane is absent;
 *  - row y+-1 on that plane: same cquad as the centre row, or across the wrap
 *    (:540-640); if (x,y+-1) is absent the whole column is absent;
 *  - node x+-1 on that row: same nquad, or across the wrap (:466-505).
 */
static void tsc(grid_t *g, int64_t s, int64_t out[27])
{
  int64_t x, y, z, L = g->L;
  node_xyz(g, s, &x, &y, &z);
  int64_t rs = g->nq_rw.v[g->nd_nq[s]], ps = g->rw_pl.v[rs], pq = g->pl_pq.v[ps];
  for (int i = 0; i < 27; i++) out[i] = -1;
  for (int dz = -1; dz <= 1; dz++) {
    int64_t p = -1;
    if (dz == 0) p = ps;
    else if (dz < 0) {
      if (z > g->pq_z0.v[pq]) p = ps - 1;
      else if (z == 0) p = find_plane(g, L - 1);
    } else {
      if (z < g->pq_z0.v[pq] + g->pq_len.v[pq] - 1) p = ps + 1;
      else if (z == L - 1 && g->pq_z0.v[0] == 0) p = g->pq_pl0.v[0];
    }
    if (p < 0) continue;
    int64_t rc = find_row(g, p, y);
    if (rc < 0 || find_node(g, rc, x) < 0) continue;
    int64_t cq = g->rw_cq.v[rc];
    for (int dy = -1; dy <= 1; dy++) {
      int64_t r = -1;
      if (dy == 0) r = rc;
      else if (dy < 0) {
        if (y > g->cq_y0.v[cq]) r = rc - 1;
        else if (y == 0) r = find_row(g, p, L - 1);
      } else {
        if (y < g->cq_y0.v[cq] + g->cq_len.v[cq] - 1) r = rc + 1;
        else if (y == L - 1) {
          int64_t c1 = g->pl_cq0.v[p];
          if (g->cq_y0.v[c1] == 0) r = g->cq_rw0.v[c1];
        }
      }
      if (r < 0) continue;
      int64_t cc = find_node(g, r, x);
      if (cc < 0) continue;
      int64_t q = g->nd_nq[cc];
      int base = (dz + 1) * 9 + (dy + 1) * 3;
      out[base + 1] = cc;
      /* x+1 */
      if (x < g->nq_x0.v[q] + g->nq_len.v[q] - 1) out[base + 2] = cc + 1;
      else if (x == L - 1) {
        int64_t f = g->rw_nq0.v[r];
        if (g->nq_x0.v[f] == 0) {
          if (g->nq_len.v[f] > 0) out[base + 2] = g->nq_nd0.v[f];
          else g->anomalies++; /* AHF would dereference an empty run */
        }
      }
      /* x-1 */
      if (x > g->nq_x0.v[q]) out[base + 0] = cc - 1;
      else if (x == 0) {
        int64_t l = g->rw_nq0.v[r + 1] - 1;
        if (g->nq_x0.v[l] + g->nq_len.v[l] == L) out[base + 0] = g->nq_nd0.v[l] + g->nq_len.v[l] - 1;
      }
    }
  }
}

static int tsc_full(const int64_t t[27])
{
  for (int i = 0; i < 27; i++)
    if (t[i] < 0) return 0;
  return 1;
}

/* f1mod, src/libutility/specific.c:120-129 */
static double f1mod(double x)
{
  if (x >= 2.0) return x - 2.0;
  if (x >= 1.0) return x - 1.0;
  return x;
}

/* TSC weights, src/libamr_serial/density.c:356-360 */
void oracle_tsc_weights(double s, double w[3])
{
  w[1] = 0.75 - s * s;
  w[0] = ((0.5 - s) * (0.5 - s)) / 2;
  w[2] = ((0.5 + s) * (0.5 + s)) / 2;
}

/* zero_dens, src/libamr_serial/density.c:443-492 */
static void zero_dens(grid_t *g)
{
  for (int64_t i = 0; i < g->nnode; i++) g->dens[i] = -1.0f;
}

/*
 * assign_npart, src/libamr_serial/density.c:238-438: nodes in traversal
 * order, particles in list order, the 27 adds k (z) outer, i (x) inner.
 * Returns FALSE when the grid has fewer than MIN_NNODES nodes (np_limit off).
 */
static int assign_npart(grid_t *g, const float *pos, const int64_t *next)
{
  uint64_t no_part = 0;
  double L = (double)g->L, shift = 0.5 / L;
  int64_t t[27];
  for (int64_t n = 0; n < g->nnode; n++) {
    int64_t x, y, z;
    node_xyz(g, n, &x, &y, &z);
    double c[3] = {f1mod(((double)x / L) + shift + 1.0), f1mod(((double)y / L) + shift + 1.0),
                   f1mod(((double)z / L) + shift + 1.0)};
    tsc(g, n, t);
    for (int64_t p = g->head[n]; p >= 0; p = next[p]) {
      no_part++;
      double w[3][3];
      for (int d = 0; d < 3; d++) {
        double s = ((double)pos[3 * p + d] - c[d]) * L;
        if (fabs(s) > 0.5 * L) {
          double a = f1mod(((double)pos[3 * p + d] + 0.5) + 1.0);
          double b = f1mod((c[d] + 0.5) + 1.0);
          s = (a - b) * L;
        }
        double ww[3];
        oracle_tsc_weights(s, ww);
        w[0][d] = ww[0]; w[1][d] = ww[1]; w[2][d] = ww[2];
      }
      for (int k = 0; k < 3; k++)
        for (int j = 0; j < 3; j++)
          for (int i = 0; i < 3; i++) {
            int64_t m = t[k * 9 + j * 3 + i];
            if (m >= 0) g->dens[m] = (float)((double)g->dens[m] + g->masstopartdens * w[k][2] * w[j][1] * w[i][0]);
          }
    }
  }
  g->size_no_part = no_part;
  g->size_no_nodes = (uint64_t)g->nnode;
  return g->nnode >= MIN_NNODES;
}

/* test_node, src/libamr_serial/refine_grid.c:113-138 */
static int test_node(grid_t *g, int64_t n)
{
  int64_t t[27];
  tsc(g, n, t);
  if (!tsc_full(t)) return 0;
  double thr = g->critdens - 1.0;
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 1; i < 3; i++)
        if ((double)g->dens[t[k * 9 + j * 3 + i]] >= thr) return 1;
  return 0;
}

static int interior(grid_t *g, int64_t n)
{
  int64_t t[27];
  tsc(g, n, t);
  return tsc_full(t);
}

/* ------------------------------------------------------- refinement ---- */

/* runs of one fine row (fine x start, length) */
typedef struct { vec64 start, len; } runs_t;

/*
 * ref_nquad over the nquad chain of coarse row r,
 * src/libamr_serial/refine_grid.c:155-303.  Appends the fine runs to out;
 * returns TRUE if any run was started (refever || refined).
 */
static int ref_row(grid_t *g, int64_t r, runs_t *out)
{
  int64_t L = g->L, q0 = g->rw_nq0.v[r], q1 = g->rw_nq0.v[r + 1];
  int any = 0;
  int64_t last = q1 - 1;
  for (int64_t q = q0; q < q1; q++) {
    int64_t x0 = g->nq_x0.v[q], len = g->nq_len.v[q], nd0 = g->nq_nd0.v[q];
    int state = 0;
    int64_t start = 0, xlen = 0, xoff = 1;
    if (x0 == 0 && g->nq_x0.v[last] + g->nq_len.v[last] == L) {
      /* periodic counterpart: test the last node of the last run first */
      if (test_node(g, g->nq_nd0.v[last] + g->nq_len.v[last] - 1)) {
        state = 1; any = 1; start = 0;
      }
      xoff = 0;
    }
    int64_t i;
    for (i = xoff; i < len - 1; i++) {
      int64_t x = x0 + i;
      if (test_node(g, nd0 + i)) {
        if (!state) { state = 1; any = 1; start = 2 * x; }
        xlen += 2;
      } else if (state) {
        state = 0;
        if (interior(g, nd0 + i)) xlen += 2; /* ghost nodes */
        v_push(&out->start, start); v_push(&out->len, xlen);
        xlen = 0;
      }
    }
    if (i < xoff) i = xoff; /* loop exit value when the loop did not run */
    if (q < last) {
      if (state) { v_push(&out->start, start); v_push(&out->len, xlen); }
    } else if (x0 + len == L && i < len && test_node(g, nd0 + i)) {
      if (!state) { any = 1; start = 2 * (x0 + i); }
      xlen += 2;
      v_push(&out->start, start); v_push(&out->len, xlen);
    } else if (x0 + len == L && i >= len) {
      g->anomalies++; /* AHF would test a node beyond the run */
      if (state) { v_push(&out->start, start); v_push(&out->len, xlen); }
    } else if (state) {
      v_push(&out->start, start); v_push(&out->len, xlen);
    }
  }
  return any;
}

/* refined coarse row: its y and fine runs */
typedef struct { int64_t y; runs_t runs; } rowres_t;
/* fine cquad: list of refined coarse rows */
typedef struct { rowres_t *rows; int64_t n, cap; } cqres_t;
typedef struct { cqres_t *cqs; int64_t n, cap; } plres_t;

static rowres_t *cq_addrow(cqres_t *c)
{
  if (c->n == c->cap) {
    c->cap = c->cap ? 2 * c->cap : 8;
    c->rows = (rowres_t *)realloc(c->rows, (size_t)c->cap * sizeof(rowres_t));
  }
  memset(&c->rows[c->n], 0, sizeof(rowres_t));
  return &c->rows[c->n++];
}
static cqres_t *pl_addcq(plres_t *p)
{
  if (p->n == p->cap) {
    p->cap = p->cap ? 2 * p->cap : 4;
    p->cqs = (cqres_t *)realloc(p->cqs, (size_t)p->cap * sizeof(cqres_t));
  }
  memset(&p->cqs[p->n], 0, sizeof(cqres_t));
  return &p->cqs[p->n++];
}
static void plres_free(plres_t *p)
{
  for (int64_t c = 0; c < p->n; c++) {
    for (int64_t r = 0; r < p->cqs[c].n; r++) {
      v_free(&p->cqs[c].rows[r].runs.start);
      v_free(&p->cqs[c].rows[r].runs.len);
    }
    free(p->cqs[c].rows);
  }
  free(p->cqs);
  memset(p, 0, sizeof(*p));
}

/*
 * tested row -> append to the current fine cquad (opening one if needed).
 * Returns whether the row refined.
 */
static int try_row(grid_t *g, int64_t r, int64_t y, plres_t *out, int *state)
{
  runs_t tmp;
  memset(&tmp, 0, sizeof(tmp));
  if (!ref_row(g, r, &tmp)) {
    v_free(&tmp.start); v_free(&tmp.len);
    return 0;
  }
  if (!*state) { pl_addcq(out); *state = 1; }
  rowres_t *rr = cq_addrow(&out->cqs[out->n - 1]);
  rr->y = y;
  rr->runs = tmp;
  return 1;
}

/*
 * ref_cquad over the cquad chain of coarse plane p,
 * src/libamr_serial/refine_grid.c:311-609.  The periodic test of
 * :357-397 looks at the first plane of the plane's pquad.
 */
static int ref_plane(grid_t *g, int64_t p, plres_t *out)
{
  int64_t L = g->L;
  int64_t pq = g->pl_pq.v[p], fpl = g->pq_pl0.v[pq];
  int64_t fc = g->pl_cq0.v[fpl], lc = g->pl_cq0.v[fpl + 1] - 1;
  int wrap = g->cq_y0.v[fc] == 0 && g->cq_y0.v[lc] + g->cq_len.v[lc] == L;
  int any = 0;
  int64_t c0 = g->pl_cq0.v[p], c1 = g->pl_cq0.v[p + 1];
  for (int64_t c = c0; c < c1; c++) {
    int64_t y0 = g->cq_y0.v[c], len = g->cq_len.v[c], rw0 = g->cq_rw0.v[c];
    int64_t lo = 1, up = -1;
    if (wrap) {
      if (y0 == 0 && y0 + len == L) { lo = 0; up = 0; }
      else if (y0 == 0) { lo = 0; up = -1; }
      else if (y0 + len == L) { lo = 1; up = 0; }
    }
    int state = 0;
    int64_t i;
    for (i = lo; i < len - 1 + up; i++) {
      if (try_row(g, rw0 + i, y0 + i, out, &state)) any = 1;
      else state = 0;
    }
    if (i < lo) i = lo;
    if (c < c1 - 1) {
      /* CASE 1: conclude, next cquad starts afresh */
    } else if (y0 + len == L) {
      /* CASE 2: one more tested row */
      if (i < len) {
        if (try_row(g, rw0 + i, y0 + i, out, &state)) any = 1;
      } else {
        g->anomalies++;
      }
    }
  }
  return any;
}

/*
 * ref_pquad over the grid's pquad chain, src/libamr_serial/refine_grid.c:617-896.
 * Result: fine pquads, each a list of refined coarse planes.
 */
typedef struct { int64_t z; plres_t pl; } plane_ref_t;
typedef struct { plane_ref_t *pl; int64_t n, cap; } pqres_t;
typedef struct { pqres_t *pq; int64_t n, cap; } gridres_t;

static void gr_free(gridres_t *r)
{
  for (int64_t i = 0; i < r->n; i++) {
    for (int64_t j = 0; j < r->pq[i].n; j++) plres_free(&r->pq[i].pl[j].pl);
    free(r->pq[i].pl);
  }
  free(r->pq);
  memset(r, 0, sizeof(*r));
}

static int try_plane(grid_t *g, int64_t p, int64_t z, gridres_t *out, int *state)
{
  plres_t tmp;
  memset(&tmp, 0, sizeof(tmp));
  if (!ref_plane(g, p, &tmp)) {
    plres_free(&tmp);
    return 0;
  }
  if (!*state) {
    if (out->n == out->cap) {
      out->cap = out->cap ? 2 * out->cap : 4;
      out->pq = (pqres_t *)realloc(out->pq, (size_t)out->cap * sizeof(pqres_t));
    }
    memset(&out->pq[out->n++], 0, sizeof(pqres_t));
    *state = 1;
  }
  pqres_t *q = &out->pq[out->n - 1];
  if (q->n == q->cap) {
    q->cap = q->cap ? 2 * q->cap : 8;
    q->pl = (plane_ref_t *)realloc(q->pl, (size_t)q->cap * sizeof(plane_ref_t));
  }
  q->pl[q->n].z = z;
  q->pl[q->n].pl = tmp;
  q->n++;
  return 1;
}

/* refine_grid, src/libamr_serial/refine_grid.c:903-985: returns TRUE and the
 * fine grid's blocks when any node was refined */
static int refine_grid(grid_t *g, grid_t *fin)
{
  int64_t L = g->L, npq = NPQ(g);
  int wrap = g->pq_z0.v[0] == 0 && g->pq_z0.v[npq - 1] + g->pq_len.v[npq - 1] == L;
  gridres_t res;
  memset(&res, 0, sizeof(res));
  int any = 0;
  for (int64_t k = 0; k < npq; k++) {
    int64_t z0 = g->pq_z0.v[k], len = g->pq_len.v[k], pl0 = g->pq_pl0.v[k];
    int64_t lo = 1, up = -1;
    if (wrap) {
      if (z0 == 0 && z0 + len == L) { lo = 0; up = 0; }
      else if (z0 == 0) { lo = 0; up = -1; }
      else if (z0 + len == L) { lo = 1; up = 0; }
    }
    int state = 0;
    int64_t i;
    for (i = lo; i < len - 1 + up; i++) {
      if (try_plane(g, pl0 + i, z0 + i, &res, &state)) any = 1;
      else state = 0;
    }
    if (i < lo) i = lo;
    if (k < npq - 1) {
      /* CASE 1 */
    } else if (z0 + len == L) {
      if (i < len) {
        if (try_plane(g, pl0 + i, z0 + i, &res, &state)) any = 1;
      } else {
        g->anomalies++;
      }
    }
  }
  if (!any) {
    gr_free(&res);
    return 0;
  }
  /* build the fine grid: every refined coarse plane/row becomes two mirrored
   * fine planes/rows (mirror_cquad/mirror_nquad, refine_grid.c:44-107) */
  memset(fin, 0, sizeof(*fin));
  fin->L = 2 * L;
  for (int64_t k = 0; k < res.n; k++) {
    pqres_t *q = &res.pq[k];
    b_pquad(fin, 2 * q->pl[0].z);
    for (int64_t j = 0; j < q->n; j++) {
      plres_t *pr = &q->pl[j].pl;
      for (int cz = 0; cz < 2; cz++) {
        b_plane(fin, 2 * q->pl[j].z + cz);
        for (int64_t c = 0; c < pr->n; c++) {
          cqres_t *cr = &pr->cqs[c];
          b_cquad(fin, 2 * cr->rows[0].y);
          for (int64_t r = 0; r < cr->n; r++)
            for (int cy = 0; cy < 2; cy++) {
              b_row(fin, 2 * cr->rows[r].y + cy);
              runs_t *ru = &cr->rows[r].runs;
              for (int64_t u = 0; u < ru->start.n; u++) b_nquad(fin, ru->start.v[u], ru->len.v[u]);
            }
        }
      }
    }
  }
  b_finish(fin);
  gr_free(&res);
  return 1;
}

/* -------------------------------------------------------------- relink ---- */

/*
 * relink, src/libamr_serial/relink.c:31-288: fine nodes in traversal order;
 * an interior fine node takes, from its mother's list and in list order, every
 * particle inside its closed cell (|x - c| <= h in all dimensions), head-
 * inserting it into its own list.
 */
static int relink(grid_t *coa, grid_t *fin, const float *pos, int64_t *next)
{
  double Lf = (double)fin->L, h = 0.5 / Lf;
  uint64_t nfp = 0;
  int64_t t[27];
  for (int64_t f = 0; f < fin->nnode; f++) {
    int64_t x, y, z;
    node_xyz(fin, f, &x, &y, &z);
    int64_t pp = find_plane(coa, z / 2);
    int64_t rr = pp < 0 ? -1 : find_row(coa, pp, y / 2);
    int64_t m = rr < 0 ? -1 : find_node(coa, rr, x / 2);
    if (m < 0) { fin->anomalies++; continue; }
    if (coa->head[m] < 0) continue;
    tsc(fin, f, t);
    if (!tsc_full(t)) continue;
    double cx = ((double)x / Lf) + h, cy = ((double)y / Lf) + h, cz = ((double)z / Lf) + h;
    int64_t prev = -1;
    for (int64_t p = coa->head[m]; p >= 0;) {
      int64_t nx = next[p];
      double dx = fabs((double)pos[3 * p] - cx), dy = fabs((double)pos[3 * p + 1] - cy),
             dz = fabs((double)pos[3 * p + 2] - cz);
      if (dx <= h && dy <= h && dz <= h) {
        if (prev < 0) coa->head[m] = nx; else next[prev] = nx;
        next[p] = fin->head[f];
        fin->head[f] = p;
        nfp++;
      } else {
        prev = p;
      }
      p = nx;
    }
  }
  fin->size_no_nodes = (uint64_t)fin->nnode;
  fin->size_no_part = nfp;
  return fin->nnode >= MIN_NNODES;
}

/* relink_back, src/libamr_serial/relink.c:294-423: append every fine list to
 * its mother's list, fine nodes in traversal order */
static void relink_back(grid_t *coa, grid_t *fin, int64_t *next)
{
  for (int64_t f = 0; f < fin->nnode; f++) {
    if (fin->head[f] < 0) continue;
    int64_t x, y, z;
    node_xyz(fin, f, &x, &y, &z);
    int64_t pp = find_plane(coa, z / 2);
    int64_t rr = pp < 0 ? -1 : find_row(coa, pp, y / 2);
    int64_t m = rr < 0 ? -1 : find_node(coa, rr, x / 2);
    if (m < 0) continue;
    if (coa->head[m] < 0) coa->head[m] = fin->head[f];
    else {
      int64_t p = coa->head[m];
      while (next[p] >= 0) p = next[p];
      next[p] = fin->head[f];
    }
    fin->head[f] = -1;
  }
}

/* ------------------------------------------------------------ driver ---- */

typedef struct {
  grid_t *g;
  int n, cap;
} gridlist_t;

static grid_t *gl_add(gridlist_t *gl)
{
  if (gl->n == gl->cap) {
    gl->cap = gl->cap ? 2 * gl->cap : 16;
    gl->g = (grid_t *)realloc(gl->g, (size_t)gl->cap * sizeof(grid_t));
  }
  memset(&gl->g[gl->n], 0, sizeof(grid_t));
  return &gl->g[gl->n++];
}

/*
 * gen_domgrids + ll + assign_npart + gen_AMRhierarchy
 * (src/libamr_serial/generate_grids.c:24-116, 126-273, 278-432;
 *  src/libamr_serial/lltools.c:32-92; src/main.c:616-648).
 */
static int64_t build_hierarchy(const snap_data *s, gridlist_t *gl, int64_t *next)
{
  int64_t N = (int64_t)s->h.npart, L = s->h.L_dom;
  const float *pos = s->pos;
  grid_t *dom = gl_add(gl);
  dom->L = L;
  double dl = (double)L;
  dom->masstodens = (dl * dl * dl) / (double)N;
  dom->masstopartdens = (dl * dl * dl) / (double)N;
  dom->critdens = s->h.Nth_dom * dom->masstodens;
  b_pquad(dom, 0);
  for (int64_t z = 0; z < L; z++) {
    b_plane(dom, z);
    b_cquad(dom, 0);
    for (int64_t y = 0; y < L; y++) {
      b_row(dom, y);
      b_nquad(dom, 0, L);
    }
  }
  b_finish(dom);
  /* ll: head insertion in array order */
  for (int64_t p = 0; p < N; p++) {
    int64_t c[3];
    for (int d = 0; d < 3; d++) {
      uint64_t n = (uint64_t)(dl * (double)pos[3 * p + d]);
      if (n > (uint64_t)(L - 1)) n = 0;
      c[d] = (int64_t)n;
    }
    int64_t nd = (c[2] * L + c[1]) * L + c[0];
    next[p] = dom->head[nd];
    dom->head[nd] = p;
  }
  zero_dens(dom);
  assign_npart(dom, pos, next);

  int64_t anomalies = 0;
  for (;;) {
    grid_t *coa = &gl->g[gl->n - 1];
    if (coa->L == s->h.L_max) break;
    grid_t fin;
    if (!refine_grid(coa, &fin)) break;
    anomalies += coa->anomalies;
    fin.masstodens = coa->masstodens * CRITMULTI;
    fin.masstopartdens = coa->masstopartdens * CRITMULTI;
    fin.critdens = s->h.Nth_ref * fin.masstodens;
    relink(coa, &fin, pos, next);
    zero_dens(&fin);
    int mg = assign_npart(&fin, pos, next);
    anomalies += fin.anomalies;
    if (!mg) {
      relink_back(coa, &fin, next);
      grid_free(&fin);
      break;
    }
    grid_t *slot = gl_add(gl);
    *slot = fin;
  }
  for (int i = 0; i < gl->n; i++) anomalies += gl->g[i].anomalies;
  return anomalies;
}

static void put(FILE *f, const void *p, size_t sz, size_t n)
{
  if (n && fwrite(p, sz, n, f) != n) { fprintf(stderr, "oracle: write failed\n"); exit(3); }
}

/* same layout as oracle/ref_driver.c dump_grid() */
static void dump_grid(FILE *f, grid_t *g, const int64_t *next)
{
  uint64_t n_listed = 0;
  for (int64_t n = 0; n < g->nnode; n++)
    for (int64_t p = g->head[n]; p >= 0; p = next[p]) n_listed++;
  uint64_t hdr[3] = {(uint64_t)g->L, g->size_no_nodes, g->size_no_part};
  double dh[2] = {g->critdens, g->masstopartdens};
  uint64_t cnt[5] = {(uint64_t)NNQ(g), (uint64_t)g->nnode, n_listed, (uint64_t)NPQ(g), (uint64_t)NCQ(g)};
  put(f, hdr, 8, 3);
  put(f, dh, 8, 2);
  put(f, cnt, 8, 5);
  for (int64_t q = 0; q < NNQ(g); q++) {
    int64_t r = g->nq_rw.v[q];
    int32_t rec[4] = {(int32_t)g->pl_z.v[g->rw_pl.v[r]], (int32_t)g->rw_y.v[r], (int32_t)g->nq_x0.v[q],
                      (int32_t)g->nq_len.v[q]};
    put(f, rec, 4, 4);
  }
  for (int64_t k = 0; k < NPQ(g); k++) {
    int32_t rec[2] = {(int32_t)g->pq_z0.v[k], (int32_t)g->pq_len.v[k]};
    put(f, rec, 4, 2);
  }
  for (int64_t c = 0; c < NCQ(g); c++) {
    int32_t rec[3] = {(int32_t)g->pl_z.v[g->cq_pl.v[c]], (int32_t)g->cq_y0.v[c], (int32_t)g->cq_len.v[c]};
    put(f, rec, 4, 3);
  }
  put(f, g->dens, 4, (size_t)g->nnode);
  for (int64_t n = 0; n < g->nnode; n++) {
    uint32_t c = 0;
    for (int64_t p = g->head[n]; p >= 0; p = next[p]) c++;
    put(f, &c, 4, 1);
  }
  for (int64_t n = 0; n < g->nnode; n++)
    for (int64_t p = g->head[n]; p >= 0; p = next[p]) {
      uint64_t v = (uint64_t)p;
      put(f, &v, 8, 1);
    }
}

int oracle_amr_dump(const char *snap_path, const char *out_path)
{
  snap_data s;
  if (snap_read(snap_path, &s)) return -1;
  int64_t *next = (int64_t *)malloc((size_t)(s.h.npart ? s.h.npart : 1) * sizeof(int64_t));
  gridlist_t gl;
  memset(&gl, 0, sizeof(gl));
  int64_t anomalies = build_hierarchy(&s, &gl, next);
  FILE *f = fopen(out_path, "wb");
  if (!f) return -2;
  put(f, "AHFGRID1", 1, 8);
  int32_t ng[2] = {gl.n, 0};
  put(f, ng, 4, 2);
  for (int i = 0; i < gl.n; i++) dump_grid(f, &gl.g[i], next);
  fclose(f);
  int n = gl.n;
  for (int i = 0; i < gl.n; i++) grid_free(&gl.g[i]);
  free(gl.g);
  free(next);
  snap_free(&s);
  if (anomalies) {
    fprintf(stderr, "oracle: %lld undefined-behaviour cases of the reference met\n", (long long)anomalies);
    return -100;
  }
  return n;
}

/* ----------------------------------------------------------- unbinding ---- */

/*
 * indexx, src/libutility/general.c:1093-1153: the Numerical Recipes
 * quicksort-index (median of three, straight insertion below 7 elements,
 * explicit stack).  arr1 and indx1 are 1-based (element 0 unused).  The
 * same partition steps are kept so that equal keys come out in the same
 * order as in AHF.
 */
void oracle_indexx(uint64_t n, const double *arr1, uint64_t *indx1)
{
  enum { SMALL = 7, STACK = 64 };
  int64_t stack[STACK + 2];
  int top = 0;
  uint64_t lo = 1, hi = n;
  for (uint64_t j = 1; j <= n; j++) indx1[j] = j;
  if (n == 0) return;
#define SWAPI(a, b) do { uint64_t t_ = (a); (a) = (b); (b) = t_; } while (0)
  for (;;) {
    if (hi - lo < SMALL) {
      for (uint64_t j = lo + 1; j <= hi; j++) {
        uint64_t it = indx1[j];
        double a = arr1[it];
        uint64_t i = j - 1;
        for (; i >= 1; i--) {
          if (arr1[indx1[i]] <= a) break;
          indx1[i + 1] = indx1[i];
        }
        indx1[i + 1] = it;
      }
      if (top == 0) break;
      hi = (uint64_t)stack[top--];
      lo = (uint64_t)stack[top--];
    } else {
      uint64_t k = (lo + hi) >> 1;
      SWAPI(indx1[k], indx1[lo + 1]);
      if (arr1[indx1[lo + 1]] > arr1[indx1[hi]]) SWAPI(indx1[lo + 1], indx1[hi]);
      if (arr1[indx1[lo]] > arr1[indx1[hi]]) SWAPI(indx1[lo], indx1[hi]);
      if (arr1[indx1[lo + 1]] > arr1[indx1[lo]]) SWAPI(indx1[lo + 1], indx1[lo]);
      uint64_t i = lo + 1, j = hi, it = indx1[lo];
      double a = arr1[it];
      for (;;) {
        do i++; while (arr1[indx1[i]] < a);
        do j--; while (arr1[indx1[j]] > a);
        if (j < i) break;
        SWAPI(indx1[i], indx1[j]);
      }
      indx1[lo] = indx1[j];
      indx1[j] = it;
      top += 2;
      if (top > STACK) { fprintf(stderr, "oracle: indexx stack overflow\n"); exit(3); }
      if (hi - i + 1 >= j - lo) {
        stack[top] = (int64_t)hi;
        stack[top - 1] = (int64_t)i;
        hi = j - 1;
      } else {
        stack[top] = (int64_t)(j - 1);
        stack[top - 1] = (int64_t)lo;
        lo = i;
      }
    }
  }
#undef SWAPI
}

typedef struct { double key; uint64_t pos; } keypos;
static int cmp_keypos(const void *a, const void *b)
{
  const keypos *x = (const keypos *)a, *y = (const keypos *)b;
  if (x->key < y->key) return -1;
  if (x->key > y->key) return 1;
  return x->pos < y->pos ? -1 : (x->pos > y->pos);
}

/* periodic separation (signed), src/libahf/ahf_halos.c:5827-5837 */
static inline double wrap_signed(double d)
{
  if (d > 0.5) d -= 1.0;
  if (d < -0.5) d += 1.0;
  return d;
}

/* sort_halo_particles, src/libahf/ahf_halos.c:5786-5894 */
void oracle_sort_halo(const float *pos, uint64_t n, uint64_t *ipart, const double c[3],
                      int32_t minpart, int stable_ties)
{
  if (n < (uint64_t)minpart) return;
  if (stable_ties) {
    keypos *kp = (keypos *)malloc(n * sizeof(keypos));
    for (uint64_t i = 0; i < n; i++) {
      const float *p = pos + 3 * ipart[i];
      double dx = wrap_signed((double)p[0] - c[0]), dy = wrap_signed((double)p[1] - c[1]),
             dz = wrap_signed((double)p[2] - c[2]);
      kp[i].key = dx * dx + dy * dy + dz * dz;
      kp[i].pos = i;
    }
    qsort(kp, n, sizeof(keypos), cmp_keypos);
    uint64_t *tmp = (uint64_t *)malloc(n * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++) tmp[i] = ipart[kp[i].pos];
    memcpy(ipart, tmp, n * sizeof(uint64_t));
    free(tmp); free(kp);
    return;
  }
  double *d = (double *)malloc((n + 1) * sizeof(double));
  uint64_t *idx = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
  uint64_t *tmp = (uint64_t *)malloc(n * sizeof(uint64_t));
  for (uint64_t i = 0; i < n; i++) {
    const float *p = pos + 3 * ipart[i];
    double dx = wrap_signed((double)p[0] - c[0]), dy = wrap_signed((double)p[1] - c[1]),
           dz = wrap_signed((double)p[2] - c[2]);
    d[i + 1] = dx * dx + dy * dy + dz * dz;
  }
  oracle_indexx(n, d, idx);
  for (uint64_t i = 0; i < n; i++) tmp[i] = ipart[idx[i + 1] - 1];
  memcpy(ipart, tmp, n * sizeof(uint64_t));
  free(d); free(idx); free(tmp);
}

/*
 * rem_unbound, src/libahf/ahf_halos.c:3292-3607 (equal particle masses,
 * Hubble drag on, no gas).  ipart is compacted in place to the bound set in
 * its original order; res receives npart, M_vir, R_vir, Phi0.
 */
void oracle_rem_unbound(const float *pos, const float *mom, uint64_t *ipart, uint64_t n,
                        const double c[3], const oracle_unbind_params *P, oracle_unbind_result *res)
{
  memset(res, 0, sizeof(*res));
  res->npart = n;
  if (n < (uint64_t)P->minpart) return;
  double v2_tune = P->vtune * P->vtune;
  uint64_t no_vbulk = (uint64_t)(P->minpart / 2);
  uint64_t nremove = 4;
  int niter = 0;
  double R_vir = 0.0;
  while (nremove > 3) {
    niter++;
    /* Phi0: mid-point integral of M(<r)/r^2 from the centre out (:3349-3427) */
    double I_now = 0.0, I_prev = 0.0, d_prev = 0.0, M_r = 0.0, Phi0 = 0.0, dist = 0.0;
    for (uint64_t j = 0; j < n; j++) {
      const float *p = pos + 3 * ipart[j];
      M_r += 1.0;
      double dX = fabs((double)p[0] - c[0]), dY = fabs((double)p[1] - c[1]), dZ = fabs((double)p[2] - c[2]);
      if (dX > 0.5) dX -= 1.0;
      if (dY > 0.5) dY -= 1.0;
      if (dZ > 0.5) dZ -= 1.0;
      dist = sqrt(dX * dX + dY * dY + dZ * dZ);
      if (dist > MACHINE_ZERO) {
        I_now = M_r / (dist * dist);
        double I_mid = (I_now + I_prev) / 2.;
        Phi0 += I_mid * (dist - d_prev);
      }
      d_prev = dist;
      I_prev = I_now;
    }
    Phi0 += M_r / dist;
    res->Phi0 = Phi0;

    /* bulk-velocity seed (:3441-3470) */
    uint64_t js;
    if (niter == 1) {
      double *m2 = (double *)calloc(no_vbulk + 1, sizeof(double));
      uint64_t *ix = (uint64_t *)calloc(no_vbulk + 1, sizeof(uint64_t));
      for (uint64_t j = 0; j < no_vbulk; j++) {
        const float *v = mom + 3 * ipart[j];
        float v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; /* FP32, as in AHF */
        m2[j + 1] = (double)v2;
      }
      oracle_indexx(no_vbulk, m2, ix);
      js = ix[(int)(no_vbulk / 2)] - 1;
      free(m2); free(ix);
    } else {
      js = 0;
    }
    const float *vs = mom + 3 * ipart[js];
    double M_vel = 1.0, VX = 1.0 * vs[0], VY = 1.0 * vs[1], VZ = 1.0 * vs[2];

    /* potential, escape test and compaction (:3472-3583) */
    nremove = 0;
    I_now = 0.0; I_prev = 0.0; d_prev = 0.0; M_r = 0.0;
    double M_vir = 0.0, Phi = 0.0;
    uint64_t nb = 0;
    for (uint64_t j = 0; j < n; j++) {
      uint64_t ip = ipart[j];
      const float *p = pos + 3 * ip, *v = mom + 3 * ip;
      M_r += 1.0;
      double dX = wrap_signed((double)p[0] - c[0]), dY = wrap_signed((double)p[1] - c[1]),
             dZ = wrap_signed((double)p[2] - c[2]);
      dist = sqrt(dX * dX + dY * dY + dZ * dZ);
      double v_esc2;
      if (dist > MACHINE_ZERO) {
        I_now = M_r / (dist * dist);
        double I_mid = (I_now + I_prev) / 2.;
        Phi += I_mid * (dist - d_prev);
        v_esc2 = (2 * fabs(Phi - Phi0) * P->phi_fac);
      } else {
        v_esc2 = 1e30;
      }
      double dVX = ((double)v[0] - VX / M_vel) * P->v_fac + P->hubble * (dX) * P->r_fac;
      double dVY = ((double)v[1] - VY / M_vel) * P->v_fac + P->hubble * (dY) * P->r_fac;
      double dVZ = ((double)v[2] - VZ / M_vel) * P->v_fac + P->hubble * (dZ) * P->r_fac;
      double vel2 = dVX * dVX + dVY * dVY + dVZ * dVZ;
      if (vel2 > v2_tune * v_esc2) {
        nremove++;
      } else {
        M_vel += 1.0;
        VX += 1.0 * v[0];
        VY += 1.0 * v[1];
        VZ += 1.0 * v[2];
        ipart[nb++] = ip; /* nb <= j: in-place compaction keeps the order */
        M_vir += 1.0;
        R_vir = dist;
        res->rvir_set = 1;
      }
      I_prev = I_now;
      d_prev = dist;
    }
    n = nb;
    res->npart = n;
    res->M_vir = M_vir;
    res->R_vir = R_vir;
    res->niter = niter;
    if ((double)n < (double)P->minpart) break;
  }
}

/*
 * rem_outsideRvir, src/libahf/ahf_halos.c:3687-3889, with AHFprofilerise
 * off (src/define.h:74) and equal masses; both calls (icall 0 and 1) cut
 * the same way, icall 0 also clears Phi0 (done by the caller).  Members are
 * taken in list order; after each one the overdensity of the sphere through
 * it is M/V * rho_fac / rho_vir, and the walk stops after the first member
 * whose sphere is below ovlim (that member is kept).  Below minpart the
 * function returns before touching anything: the result is n and nothing is
 * written.
 */
uint64_t oracle_rem_outsideRvir(const float *pos, const uint64_t *ipart, uint64_t n, const double c[3],
                                const oracle_rvir_params *P, double *M_vir, double *R_vir, double *ovdens)
{
  if (n < (uint64_t)P->minpart) return n;
  uint64_t kept = 0;
  double M = 0.0, R = -1.0, od = 2 * P->ovlim;
  while (kept < n && od >= P->ovlim) {
    const float *p = pos + 3 * ipart[kept];
    M += 1.0;
    double dX = fabs((double)p[0] - c[0]), dY = fabs((double)p[1] - c[1]), dZ = fabs((double)p[2] - c[2]);
    if (dX > 0.5) dX -= 1.0;
    if (dY > 0.5) dY -= 1.0;
    if (dZ > 0.5) dZ -= 1.0;
    R = sqrt(dX * dX + dY * dY + dZ * dZ);
    const double V = 4. * PI_AHF / 3. * (R * R * R);
    od = M / V * P->rho_fac / P->rho_vir;
    kept++;
  }
  *M_vir = M;
  *R_vir = R;
  *ovdens = od;
  return kept;
}

/*
 * The part of ahf_halos_sfc_constructHalo (src/libahf/ahf_halos_sfc.c:140-153)
 * between the gather and the profiles, on one halo: sort_halo_particles,
 * rem_outsideRvir(0), rem_unbound, rem_outsideRvir(1).  st starts as a
 * zeroed HALO would; each step writes only the fields the reference step
 * writes.  ipart is reordered and truncated in place.
 */
void oracle_construct_halo(const float *pos, const float *mom, uint64_t *ipart, uint64_t n, const double c[3],
                           const oracle_unbind_params *U, const oracle_rvir_params *R, int stable_ties,
                           oracle_halo_state *st)
{
  st->npart = n;
  oracle_sort_halo(pos, n, ipart, c, U->minpart, stable_ties);
  if (n >= (uint64_t)R->minpart) {
    st->npart = oracle_rem_outsideRvir(pos, ipart, n, c, R, &st->M_vir, &st->R_vir, &st->ovdens);
    st->Phi0 = 0.0;
    st->set |= ORACLE_SET_MVIR | ORACLE_SET_RVIR | ORACLE_SET_OVDENS | ORACLE_SET_PHI0;
  }
  if (st->npart >= (uint64_t)U->minpart) {
    oracle_unbind_result r;
    oracle_rem_unbound(pos, mom, ipart, st->npart, c, U, &r);
    st->npart = r.npart;
    st->M_vir = r.M_vir;
    st->Phi0 = r.Phi0;
    st->niter = r.niter;
    st->set |= ORACLE_SET_MVIR | ORACLE_SET_PHI0;
    /* AHF stores its R_vir local, which stays unassigned when no member was
     * ever bound */
    if (r.rvir_set) {
      st->R_vir = r.R_vir;
      st->set |= ORACLE_SET_RVIR;
    } else {
      st->R_vir = 0.0;
      st->set &= ~ORACLE_SET_RVIR;
    }
  }
  if (st->npart >= (uint64_t)R->minpart) {
    st->npart = oracle_rem_outsideRvir(pos, ipart, st->npart, c, R, &st->M_vir, &st->R_vir, &st->ovdens);
    st->set |= ORACLE_SET_MVIR | ORACLE_SET_RVIR | ORACLE_SET_OVDENS;
  }
}

static void dump_states(FILE *f, const snap_data *s, const uint64_t *off, const oracle_halo_state *st)
{
  const uint64_t nh = s->h.nhalos;
  put(f, "AHFHALO1", 1, 8);
  put(f, &nh, 8, 1);
  for (uint64_t h = 0; h < nh; h++) {
    uint64_t n = st[h].npart;
    double v[4] = {st[h].M_vir, st[h].R_vir, st[h].Phi0, st[h].ovdens};
    put(f, &n, 8, 1);
    put(f, v, 8, 4);
    put(f, s->ipart + off[h], 8, (size_t)n);
  }
}

int oracle_construct_dump(const char *snap_path, const char *out_path, int threads, int stable_ties)
{
  snap_data s;
  if (snap_read(snap_path, &s)) return -1;
  const uint64_t nh = s.h.nhalos;
  uint64_t *off = (uint64_t *)malloc((nh + 1) * sizeof(uint64_t));
  off[0] = 0;
  for (uint64_t i = 0; i < nh; i++) off[i + 1] = off[i] + s.halo_npart[i];
  oracle_halo_state *st = (oracle_halo_state *)calloc(nh ? nh : 1, sizeof(oracle_halo_state));
  const oracle_unbind_params U = {s.h.minpart, s.h.vtune, s.h.r_fac, s.h.v_fac, s.h.phi_fac, s.h.hubble};
  const double b = s.h.boxsize;
  const oracle_rvir_params R = {s.h.minpart, s.h.ovlim, s.h.rho_vir, s.h.pmass / (b * b * b)};
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  long i;
#pragma omp parallel for schedule(dynamic) private(i)
  for (i = 0; i < (long)nh; i++)
    oracle_construct_halo(s.pos, s.mom, s.ipart + off[i], s.halo_npart[i], s.centre + 3 * i, &U, &R, stable_ties,
                          &st[i]);
  FILE *f = fopen(out_path, "wb");
  if (!f) return -2;
  dump_states(f, &s, off, st);
  fclose(f);
  free(off); free(st);
  snap_free(&s);
  return 0;
}

int oracle_unbind_dump(const char *snap_path, const char *out_path, int threads, int stable_ties)
{
  snap_data s;
  if (snap_read(snap_path, &s)) return -1;
  uint64_t nh = s.h.nhalos;
  uint64_t *off = (uint64_t *)malloc((nh + 1) * sizeof(uint64_t));
  off[0] = 0;
  for (uint64_t i = 0; i < nh; i++) off[i + 1] = off[i] + s.halo_npart[i];
  oracle_unbind_result *res = (oracle_unbind_result *)calloc(nh ? nh : 1, sizeof(oracle_unbind_result));
  oracle_unbind_params P = {s.h.minpart, s.h.vtune, s.h.r_fac, s.h.v_fac, s.h.phi_fac, s.h.hubble};
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  long i;
#pragma omp parallel for schedule(dynamic) private(i)
  for (i = 0; i < (long)nh; i++) {
    uint64_t *ip = s.ipart + off[i];
    oracle_sort_halo(s.pos, s.halo_npart[i], ip, s.centre + 3 * i, s.h.minpart, stable_ties);
    oracle_rem_unbound(s.pos, s.mom, ip, s.halo_npart[i], s.centre + 3 * i, &P, &res[i]);
  }
  FILE *f = fopen(out_path, "wb");
  if (!f) return -2;
  put(f, "AHFHALO1", 1, 8);
  put(f, &nh, 8, 1);
  for (uint64_t h = 0; h < nh; h++) {
    uint64_t n = res[h].npart;
    double v[4] = {res[h].M_vir, res[h].R_vir, res[h].Phi0, 0.0};
    put(f, &n, 8, 1);
    put(f, v, 8, 4);
    put(f, s.ipart + off[h], 8, (size_t)n);
  }
  fclose(f);
  free(off); free(res);
  snap_free(&s);
  return 0;
}

/* ===================================================== SFC keys, gather */

#define SFC_BITS 21          /* BITS_PER_DIMENSION, src/param.h:92 */
#define GATHER_FAC 1.001     /* GATHERRAD_FAC, src/libahf/ahf_halos_sfc.c:46 */

/*
 * Hilbert index of grid cell c[0..2] on a 2^bits grid: the curve of
 * src/libsfc/hilbert.c:197-241 (Butz's algorithm as D. Moore implements
 * it), restated digit by digit.  The per-level digits hold one bit of each
 * coordinate (x in bit 0); each is XORed with the digit of the level above,
 * then, from the top, with a flip bit and rotated right by a rotation that
 * both depend on the previous output digit.  The output digits below the
 * top one get bit 2 flipped, and the whole index is Gray-decoded.
 */
uint64_t oracle_hilbert_c2i(const uint64_t c[3], unsigned bits)
{
  unsigned dig[64];
  for (unsigned l = 0; l < bits; l++) {
    const unsigned b = bits - 1 - l;
    dig[l] = (unsigned)(((c[0] >> b) & 1) | (((c[1] >> b) & 1) << 1) | (((c[2] >> b) & 1) << 2));
  }
  for (unsigned l = bits; l-- > 1;) dig[l] ^= dig[l - 1];
  uint64_t h = 0;
  unsigned rot = 0, flip = 0;
  for (unsigned l = 0; l < bits; l++) {
    const unsigned v = (dig[l] ^ flip) & 7u;
    const unsigned t = rot ? ((v >> rot) | (v << (3 - rot))) & 7u : v;
    h = (h << 3) | t;
    flip = 1u << rot;
    rot = (rot + 1 + ((t & 1u) ? 1 : (t & 2u) ? 2 : 0)) % 3;
  }
  for (unsigned l = 1; l < bits; l++) h ^= (uint64_t)4 << (3 * (bits - 1 - l));
  uint64_t r = 0, acc = 0;
  for (int i = 3 * (int)bits - 1; i >= 0; i--) {
    acc ^= (h >> i) & 1;
    r |= acc << i;
  }
  return r;
}

/* sfc_curve_calcKey (src/libsfc/hilbert_util.c:68-92): truncate to the
 * 2^bits grid, the last cell taking the upper face */
uint64_t oracle_sfc_key(double x, double y, double z, unsigned bits)
{
  const uint64_t max = (uint64_t)1 << bits;
  const double p[3] = {x, y, z};
  uint64_t c[3];
  for (int d = 0; d < 3; d++) {
    c[d] = (uint64_t)trunc(p[d] * (double)max);
    if (c[d] == max) c[d] = max - 1;
  }
  return oracle_hilbert_c2i(c, bits);
}

/* first index of the sorted keys at or above k (what util_bsearch followed
 * by the step back over equal keys returns in local_findMinOffset,
 * src/libahf/ahf_halos_sfc.c:370-414), n if none */
static uint64_t lower_key(const uint64_t *key, uint64_t n, uint64_t k)
{
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (key[mid] < k) lo = mid + 1; else hi = mid;
  }
  return lo;
}

static double sep1(double a, double b)
{
  double d = fabs(a - b);
  if (d > 0.5) d = 1.0 - d;
  return d;
}

/*
 * ahf_halos_sfc_gatherParts, src/libahf/ahf_halos_sfc.c:171-367, for one
 * halo: the members (indices into the key-sorted particles) are appended
 * to out (capacity grown with realloc) in the order the reference adds them.
 * Returns the member count; *n_cand receives the particles tested.
 */
uint64_t oracle_gather_halo(const float *pos, const uint64_t *key, uint64_t n, const double c[3], double rad,
                            uint64_t **out, uint64_t *n_cand)
{
  unsigned bits = 1;
  while ((GATHER_FAC * rad < 1. / (1 << (bits + 1))) && (bits + 1) <= SFC_BITS) bits++;
  const double rad2 = rad * rad;
  uint64_t ncell = 0, cells[27];
  if (bits == 1) {
    for (uint64_t k = 0; k < 8; k++) cells[ncell++] = k;
  } else {
    const uint64_t m = (uint64_t)1 << bits;
    uint64_t b[3];
    for (int d = 0; d < 3; d++) {
      b[d] = (uint64_t)trunc(c[d] * (double)m);
      if (b[d] == m) b[d] = m - 1;
    }
    const double g = 1. / (double)m;
    const double corner = 0.5 * sqrt(3.) * g;
    for (int i = -1; i <= 1; i++)
      for (int j = -1; j <= 1; j++)
        for (int k = -1; k <= 1; k++) {
          const uint64_t q[3] = {(b[0] + m + i) % m, (b[1] + m + j) % m, (b[2] + m + k) % m};
          double d2 = 0.0;
          const double dx = sep1(g * q[0] + 0.5 * g, c[0]), dy = sep1(g * q[1] + 0.5 * g, c[1]),
                       dz = sep1(g * q[2] + 0.5 * g, c[2]);
          d2 = dx * dx + dy * dy + dz * dz;
          if (sqrt(d2) > corner + GATHER_FAC * rad) continue;
          cells[ncell++] = oracle_hilbert_c2i(q, bits);
        }
  }
  const unsigned sh = 3 * (SFC_BITS - bits);
  uint64_t cnt = 0, cap = 0, tested = 0;
  for (uint64_t i = 0; i < ncell; i++) {
    const uint64_t kmin = cells[i] << sh, kmax = kmin + (((uint64_t)1 << sh) - 1);
    for (uint64_t o = lower_key(key, n, kmin); o < n && key[o] <= kmax; o++) {
      tested++;
      const float *p = pos + 3 * o;
      const double dx = sep1((double)p[0], c[0]), dy = sep1((double)p[1], c[1]), dz = sep1((double)p[2], c[2]);
      if (dx * dx + dy * dy + dz * dz <= rad2) {
        if (cnt == cap) {
          cap = cap ? 2 * cap : 64;
          *out = (uint64_t *)realloc(*out, cap * sizeof(uint64_t));
        }
        (*out)[cnt++] = o;
      }
    }
  }
  if (n_cand) *n_cand = tested;
  return cnt;
}

/* keys of every particle of a snapshot (main.c:343-350); 1 if they are in
 * non-decreasing order, as AHF's sort leaves them */
static int snap_keys(const snap_data *s, uint64_t *key)
{
  const uint64_t n = s->h.npart;
  long i;
#pragma omp parallel for schedule(static) private(i)
  for (i = 0; i < (long)n; i++)
    key[i] = oracle_sfc_key((double)s->pos[3 * i], (double)s->pos[3 * i + 1], (double)s->pos[3 * i + 2], SFC_BITS);
  for (uint64_t j = 1; j < n; j++)
    if (key[j] < key[j - 1]) return 0;
  return 1;
}

/*
 * The whole of ahf_halos_sfc_constructHalo (src/libahf/ahf_halos_sfc.c:
 * 119-157) but the profiles, for every halo of the snapshot: haloes whose
 * snapshot list is empty are skipped (:122-123); the others get npart = 0,
 * then gatherParts from the centre and gather radius, and, when construct
 * is set, sort_halo_particles, rem_outsideRvir(0), rem_unbound,
 * rem_outsideRvir(1).  Writes the AHFHALO1 dump; *n_cand receives the
 * particles the gather tested.  Returns 0, -1 (bad snapshot) or -3
 * (particles not in key order).
 */
int oracle_gather_dump(const char *snap_path, const char *out_path, int threads, int stable_ties, int construct,
                       uint64_t *n_cand)
{
  snap_data s;
  if (snap_read(snap_path, &s)) return -1;
  const uint64_t nh = s.h.nhalos, n = s.h.npart;
  uint64_t *key = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
  if (!snap_keys(&s, key)) {
    free(key);
    snap_free(&s);
    return -3;
  }
  uint64_t **lists = (uint64_t **)calloc(nh ? nh : 1, sizeof(uint64_t *));
  uint64_t *cand = (uint64_t *)calloc(nh ? nh : 1, sizeof(uint64_t));
  oracle_halo_state *st = (oracle_halo_state *)calloc(nh ? nh : 1, sizeof(oracle_halo_state));
  const oracle_unbind_params U = {s.h.minpart, s.h.vtune, s.h.r_fac, s.h.v_fac, s.h.phi_fac, s.h.hubble};
  const double b = s.h.boxsize;
  const oracle_rvir_params R = {s.h.minpart, s.h.ovlim, s.h.rho_vir, s.h.pmass / (b * b * b)};
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  long i;
#pragma omp parallel for schedule(dynamic) private(i)
  for (i = 0; i < (long)nh; i++) {
    if (s.halo_npart[i] == 0) continue;
    const uint64_t m = oracle_gather_halo(s.pos, key, n, s.centre + 3 * i, s.gather_rad[i], &lists[i], &cand[i]);
    st[i].npart = m;
    if (construct)
      oracle_construct_halo(s.pos, s.mom, lists[i], m, s.centre + 3 * i, &U, &R, stable_ties, &st[i]);
  }
  uint64_t tot = 0;
  for (uint64_t h = 0; h < nh; h++) tot += cand[h];
  if (n_cand) *n_cand = tot;
  FILE *f = fopen(out_path, "wb");
  int rc = f ? 0 : -2;
  if (f) {
    put(f, "AHFHALO1", 1, 8);
    put(f, &nh, 8, 1);
    for (uint64_t h = 0; h < nh; h++) {
      const uint64_t m = s.halo_npart[h] == 0 ? 0 : st[h].npart;
      double v[4] = {st[h].M_vir, st[h].R_vir, st[h].Phi0, st[h].ovdens};
      put(f, &m, 8, 1);
      put(f, v, 8, 4);
      if (m) put(f, lists[h], 8, (size_t)m);
    }
    fclose(f);
  }
  for (uint64_t h = 0; h < nh; h++) free(lists[h]);
  free(lists); free(cand); free(st); free(key);
  snap_free(&s);
  return rc;
}

/* ======================================================== HaloProfiles */

/* HaloProfiles (src/libahf/ahf_halos.c:3961-5060) of one constructed halo:
 * the members ip[0..n) in their radial order, Phi0 and R_vir as the
 * construction left them.  Returns nbins (0 below minpart). */
int oracle_halo_profiles(const float *pos, const float *mom, const uint64_t *ip, uint64_t n, const double c[3],
                         double Phi0, double R_vir, const oracle_prof_params *p, double *props, double *prof)
{
  const ahf_prof_consts K = {p->minpart, p->r_fac, p->v_fac, p->m_fac, p->phi_fac, p->hubble};
  if (n < (uint64_t)p->minpart || n == 0) return 0;
  double *scratch = (double *)malloc(6 * n * sizeof(double));
  if (!scratch) return -1;
  const int nb = ahf_halo_profiles((int64_t)n, ip, pos, pos + 1, pos + 2, mom, mom + 1, mom + 2, 3, c, Phi0, R_vir,
                                   &K, scratch, props, prof);
  free(scratch);
  return nb;
}

int oracle_prof_nbins(uint64_t n)
{
  return ahf_prof_nbins((int64_t)n);
}

void oracle_prof_get_axes(double it[9], double axes[3])
{
  double t[3][3];
  for (int i = 0; i < 9; i++) t[i / 3][i % 3] = it[i];
  ahf_prof_get_axes(t, &axes[0], &axes[1], &axes[2]);
  for (int i = 0; i < 9; i++) it[i] = t[i / 3][i % 3];
}

double oracle_prof_cnfw(double V2_max, double V2_vir) { return ahf_prof_cnfw(V2_max, V2_vir); }
double oracle_prof_cr1(double R1) { return ahf_prof_cr1(R1); }

void oracle_prof_find_max(const double *x, const double *y, int64_t n, int ismooth, double *x_max, double *y_max)
{
  double *w = (double *)malloc(2 * (size_t)(n > 0 ? n : 1) * sizeof(double));
  ahf_prof_find_max(x, y, n, ismooth, w, w + n, x_max, y_max);
  free(w);
}

/* the halo construction of every halo of the snapshot (from its lists, or
 * gathered from its centre and radius with gather = 1), then HaloProfiles;
 * AHFPROF1 dump: nhalos, then per halo npart, nbins, and for nbins > 0 the
 * ORACLE_PROF_NP properties and nbins x ORACLE_PROF_NQ profile values */
int oracle_profiles_dump(const char *snap_path, const char *out_path, int threads, int gather)
{
  snap_data s;
  if (snap_read(snap_path, &s)) return -1;
  const uint64_t nh = s.h.nhalos, n = s.h.npart;
  uint64_t *key = NULL;
  if (gather) {
    key = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    if (!snap_keys(&s, key)) {
      free(key);
      snap_free(&s);
      return -3;
    }
  }
  uint64_t *off = (uint64_t *)malloc((nh + 1) * sizeof(uint64_t));
  off[0] = 0;
  for (uint64_t i = 0; i < nh; i++) off[i + 1] = off[i] + s.halo_npart[i];
  uint64_t **lists = (uint64_t **)calloc(nh ? nh : 1, sizeof(uint64_t *));
  oracle_halo_state *st = (oracle_halo_state *)calloc(nh ? nh : 1, sizeof(oracle_halo_state));
  int *nbins = (int *)calloc(nh ? nh : 1, sizeof(int));
  double *props = (double *)calloc((nh ? nh : 1) * AHF_PROF_NP, sizeof(double));
  double *prof = (double *)calloc((nh ? nh : 1) * AHF_PROF_MAXBINS * AHF_PROF_NQ, sizeof(double));
  const oracle_unbind_params U = {s.h.minpart, s.h.vtune, s.h.r_fac, s.h.v_fac, s.h.phi_fac, s.h.hubble};
  const double b = s.h.boxsize;
  const oracle_rvir_params R = {s.h.minpart, s.h.ovlim, s.h.rho_vir, s.h.pmass / (b * b * b)};
  const oracle_prof_params P = {s.h.minpart, s.h.r_fac, s.h.v_fac, s.h.pmass, s.h.phi_fac, s.h.hubble};
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  long i;
#pragma omp parallel for schedule(dynamic) private(i)
  for (i = 0; i < (long)nh; i++) {
    uint64_t m = s.halo_npart[i];
    if (gather) {
      if (m == 0) continue;
      uint64_t cand;
      m = oracle_gather_halo(s.pos, key, n, s.centre + 3 * i, s.gather_rad[i], &lists[i], &cand);
    } else {
      lists[i] = (uint64_t *)malloc((m ? m : 1) * sizeof(uint64_t));
      memcpy(lists[i], s.ipart + off[i], m * sizeof(uint64_t));
    }
    st[i].npart = m;
    oracle_construct_halo(s.pos, s.mom, lists[i], m, s.centre + 3 * i, &U, &R, 0, &st[i]);
    nbins[i] = oracle_halo_profiles(s.pos, s.mom, lists[i], st[i].npart, s.centre + 3 * i, st[i].Phi0, st[i].R_vir,
                                    &P, props + AHF_PROF_NP * i, prof + (size_t)AHF_PROF_MAXBINS * AHF_PROF_NQ * i);
  }
  FILE *f = fopen(out_path, "wb");
  int rc = f ? 0 : -2;
  if (f) {
    put(f, "AHFPROF1", 1, 8);
    put(f, &nh, 8, 1);
    for (uint64_t h = 0; h < nh; h++) {
      const uint64_t v[2] = {st[h].npart, (uint64_t)(nbins[h] > 0 ? nbins[h] : 0)};
      put(f, v, 8, 2);
      if (v[1]) {
        put(f, props + AHF_PROF_NP * h, 8, AHF_PROF_NP);
        put(f, prof + (size_t)AHF_PROF_MAXBINS * AHF_PROF_NQ * h, 8, (size_t)v[1] * AHF_PROF_NQ);
      }
    }
    fclose(f);
  }
  for (uint64_t h = 0; h < nh; h++) free(lists[h]);
  free(lists); free(st); free(nbins); free(props); free(prof); free(off); free(key);
  snap_free(&s);
  return rc;
}

/* ------------------------------------------------------------ gridinfo ---- */

/*
 * ahf_gridinfo, src/libahf/ahf_gridinfo.c:59-860: the spatially connected
 * refinements of every grid from min_ref on.  Restated on the grids of
 * build_hierarchy (nodes in traversal order):
 *   - min_ref: the last grid i (counting the domain as 0) whose refinement
 *     overdensity (Nth_ref * refine_mass / (boxsize/L)^3) / rho_vir lies
 *     below ovlim (:147-175); max_ovdens is the last grid's;
 *   - per grid, colours zeroed, then every node in traversal order looks at
 *     its six face neighbours as get_TSCnodes finds them (colourInfo,
 *     :1126-1310: -y, -x, -z, +y, +x, +z, distinct non-zero colours in
 *     that order of discovery, sorted ascending into three holders): none
 *     coloured -> a new colour and a new SPATIALREF at the list's tail; one
 *     -> it; two -> the lower, the higher colour deleted and replaced; three
 *     -> the lowest, the other two deleted and replaced.  A replacement
 *     restarts the node loop at the deleted colour's first node (the one
 *     deleted last when two are) and recolours until the first uncoloured
 *     node, which the loop then takes as usual (:206-560);
 *   - periodic flags: a node on a low face (x, y or z = 0) whose neighbour
 *     across that face exists marks its colour's SPATIALREF (:604-690,
 *     testBound :1090-1124);
 *   - spatialRefIndex[name] = {grid, rank among that grid's refinements in
 *     list order, periodic flags} (:700-760).
 * More than three distinct colours around one node overflow AHF's holder
 * (undefined behaviour): reported, as the AMR anomalies are.
 */
typedef struct {
  int32_t name, level;
  int64_t start; /* node of the grid where the colour began */
  int32_t px, py, pz;
  int alive;
} gi_ref;

typedef struct {
  int32_t holder[3];
  int num_replace;
} gi_col;

static int gi_colour_info(const int32_t *colour, const int64_t t[27], gi_col *c)
{
  /* r = -y, g = -x, b = -z, t = +y, h = +x, n = +z */
  const int nb[6] = {1 * 9 + 0 * 3 + 1, 1 * 9 + 1 * 3 + 0, 0 * 9 + 1 * 3 + 1, 1 * 9 + 2 * 3 + 1, 1 * 9 + 1 * 3 + 2,
                     2 * 9 + 1 * 3 + 1};
  int32_t seen[6];
  int ncol = 0;
  for (int i = 0; i < 6; i++) {
    const int32_t v = t[nb[i]] >= 0 ? colour[t[nb[i]]] : 0;
    if (!v) continue;
    int old = 0;
    for (int j = 0; j < ncol; j++) old |= seen[j] == v;
    if (!old) seen[ncol++] = v;
  }
  if (ncol > 3) return -1;
  c->holder[0] = c->holder[1] = c->holder[2] = 0;
  for (int i = 0; i < ncol; i++) c->holder[i] = seen[i];
  /* sort the three holders ascending (zeros first) */
  for (int i = 0; i < 3; i++)
    for (int j = i + 1; j < 3; j++)
      if (c->holder[j] < c->holder[i]) {
        const int32_t x = c->holder[i];
        c->holder[i] = c->holder[j];
        c->holder[j] = x;
      }
  const int32_t r = c->holder[0], g = c->holder[1], b = c->holder[2];
  const int nz = (r == 0) + (g == 0) + (b == 0);
  if (nz == 3) c->num_replace = 0;
  else if (nz == 2) c->num_replace = 1;
  else if (nz == 1) c->num_replace = g == b ? 1 : 2;
  else if (r == g && r == b) c->num_replace = 1;
  else if ((r == g && r != b) || (r == b && r != g)) c->num_replace = 2;
  else if (g == b && g != r) c->num_replace = 3;
  else c->num_replace = 4;
  return 0;
}

/* the replacement rule of a recolouring sweep at one node */
static void gi_replace(int32_t *colour, int64_t n, const gi_col *c)
{
  const int32_t cur = colour[n];
  if (c->num_replace == 2) {
    if (cur == c->holder[2]) colour[n] = c->holder[1];
  } else if (c->num_replace == 3) {
    if (cur == c->holder[2]) colour[n] = c->holder[0];
  } else if (c->num_replace == 4) {
    if (cur == c->holder[2]) colour[n] = c->holder[0];
    if (cur == c->holder[1]) colour[n] = c->holder[0];
  }
}

/* remove the SPATIALREF of colour `name`, returning its first node */
static int64_t gi_delete(gi_ref *refs, int64_t nref, int32_t name)
{
  for (int64_t i = 0; i < nref; i++)
    if (refs[i].alive && refs[i].name == name) {
      refs[i].alive = 0;
      return refs[i].start;
    }
  return -1;
}

int oracle_gridinfo_dump(const char *snap_path, const char *out_path, double ovlim, double rho_vir,
                         double refine_mass)
{
  snap_data s;
  if (snap_read(snap_path, &s)) return -1;
  int64_t *next = (int64_t *)malloc((size_t)(s.h.npart ? s.h.npart : 1) * sizeof(int64_t));
  gridlist_t gl;
  memset(&gl, 0, sizeof(gl));
  int64_t anomalies = build_hierarchy(&s, &gl, next);
  /* min_ref and max_ovdens (:147-175) */
  const int nref_grids = gl.n - 1;
  int start = 1;
  double refine_ovdens = 0.0;
  for (int i = 0; i <= nref_grids; i++) {
    const double fl1dim = (double)gl.g[i].L;
    const double len = s.h.boxsize / fl1dim;
    const double vol = len * len * len;
    refine_ovdens = (s.h.Nth_ref * refine_mass / vol) / rho_vir;
    if (refine_ovdens < ovlim) start = i;
  }
  const int min_ref = start, no_grids = nref_grids - min_ref + 1;
  int32_t **colour = (int32_t **)calloc((size_t)(no_grids > 0 ? no_grids : 1), sizeof(int32_t *));
  int64_t cap = 1024, nref = 0;
  gi_ref *refs = (gi_ref *)malloc((size_t)cap * sizeof(gi_ref));
  int32_t col_counter = 0;
  int replace = 0, ub = 0;
  gi_col c;
  memset(&c, 0, sizeof(c));
  int64_t t[27];
  for (int lv = 0; lv < no_grids; lv++) {
    grid_t *g = &gl.g[min_ref + lv];
    int32_t *cl = colour[lv] = (int32_t *)calloc((size_t)(g->nnode ? g->nnode : 1), sizeof(int32_t));
    for (int64_t n = 0; n < g->nnode; n++) {
      if (replace && cl[n] == 0) replace = 0;
      if (replace) {
        gi_replace(cl, n, &c);
        continue;
      }
      tsc(g, n, t);
      if (gi_colour_info(cl, t, &c)) {
        ub++;
        c.num_replace = 1;
      }
      if (c.num_replace == 0) {
        col_counter++;
        cl[n] = col_counter;
        if (nref == cap) {
          cap *= 2;
          refs = (gi_ref *)realloc(refs, (size_t)cap * sizeof(gi_ref));
        }
        refs[nref++] = (gi_ref){col_counter, lv, n, 0, 0, 0, 1};
      } else if (c.num_replace == 1) {
        cl[n] = c.holder[2];
      } else {
        int64_t s0;
        if (c.num_replace == 2) {
          cl[n] = c.holder[1];
          s0 = gi_delete(refs, nref, c.holder[2]);
        } else if (c.num_replace == 3) {
          cl[n] = c.holder[0];
          s0 = gi_delete(refs, nref, c.holder[2]);
        } else {
          cl[n] = c.holder[0];
          (void)gi_delete(refs, nref, c.holder[2]);
          s0 = gi_delete(refs, nref, c.holder[1]);
        }
        if (s0 < 0) {
          ub++;
          continue;
        }
        /* back to the deleted colour's first node; the loop goes on from
         * the node after it, recolouring until an uncoloured node */
        replace = 1;
        gi_replace(cl, s0, &c);
        n = s0;
      }
    }
  }
  /* periodic flags (:604-690) */
  for (int lv = 0; lv < no_grids; lv++) {
    grid_t *g = &gl.g[min_ref + lv];
    for (int64_t n = 0; n < g->nnode; n++) {
      int64_t x, y, z;
      node_xyz(g, n, &x, &y, &z);
      if (x != 0 && y != 0 && z != 0) continue;
      tsc(g, n, t);
      const int py = y == 0 && t[1 * 9 + 0 * 3 + 1] >= 0, px = x == 0 && t[1 * 9 + 1 * 3 + 0] >= 0,
                pz = z == 0 && t[0 * 9 + 1 * 3 + 1] >= 0;
      if (!(px || py || pz)) continue;
      for (int64_t i = 0; i < nref; i++)
        if (refs[i].alive && refs[i].name == colour[lv][n]) {
          refs[i].px |= px;
          refs[i].py |= py;
          refs[i].pz |= pz;
          break;
        }
    }
  }
  /* spatialRefIndex and numIsoRef (:700-760) */
  int32_t *num_iso = (int32_t *)calloc((size_t)(no_grids > 0 ? no_grids : 1), sizeof(int32_t));
  int32_t *iso = (int32_t *)calloc((size_t)(nref ? nref : 1), sizeof(int32_t));
  int32_t tot = 0, alive = 0;
  for (int64_t i = 0; i < nref; i++)
    if (refs[i].alive) {
      iso[i] = num_iso[refs[i].level]++;
      tot++;
      alive++;
    }
  FILE *f = fopen(out_path, "wb");
  if (!f) return -2;
  const int32_t hdr[4] = {min_ref, no_grids, tot, alive};
  put(f, "AHFGINF1", 1, 8);
  put(f, hdr, 4, 4);
  put(f, &refine_ovdens, 8, 1);
  put(f, num_iso, 4, (size_t)(no_grids > 0 ? no_grids : 0));
  for (int lv = 0; lv < no_grids; lv++) {
    const uint64_t n = (uint64_t)gl.g[min_ref + lv].nnode;
    put(f, &n, 8, 1);
    put(f, colour[lv], 4, (size_t)n);
  }
  /* surviving colours, ascending (the list is in creation order) */
  for (int64_t i = 0; i < nref; i++)
    if (refs[i].alive) {
      const int32_t e[6] = {refs[i].name, refs[i].level, iso[i], refs[i].px, refs[i].py, refs[i].pz};
      put(f, e, 4, 6);
    }
  fclose(f);
  for (int lv = 0; lv < no_grids; lv++) free(colour[lv]);
  free(colour); free(refs); free(num_iso); free(iso);
  for (int i = 0; i < gl.n; i++) grid_free(&gl.g[i]);
  free(gl.g);
  free(next);
  snap_free(&s);
  if (ub || anomalies) {
    fprintf(stderr, "oracle: %lld undefined-behaviour cases of the reference met\n", (long long)(ub + anomalies));
    return -100;
  }
  return no_grids;
}
